// api.cpp -- C ABI of the MI355X FDM engine (include/fullmag_fdm.h and
// include/fullmag_fdm_ext.h).  Host-side counterpart of the reference's
// native/backends/fdm/src/api.cpp:151-1070 and context.cu: plan validation,
// device allocation, uploads/downloads, step driving, snapshots.
//
// Step semantics follow the reference CPU engine (crates/fullmag-engine):
// the trial dt is clamped to [dt_min, dt_max], rejected trials shrink dt and
// retry until one is accepted, and the accepted step reports the
// controller's suggested next dt.  The trial/accept loop runs on the device;
// the host reads one small control block per accepted step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fullmag_fdm.h"
#include "../../include/fullmag_fdm_ext.h"
#include "engine.hpp"

namespace fm {
template <class T>
hipError_t upload_methods();
int method_nstage(int method);
int method_stage_nk(int method, int s);
int method_efuse(int method);
template <class T>
const char *configure(Args<T> &A, size_t lds_limit, int rows_req = 0, int xfast_req = -1);
template <class T>
long long num_partials(const Args<T> &A);
template <class T>
hipError_t enqueue_eval(const Args<T> &A, int s, int stats, hipStream_t st);
template <class T>
hipError_t enqueue_trial(const Args<T> &A, int nstage, hipStream_t st);
template <class T>
hipError_t enqueue_k1(const Args<T> &A, hipStream_t st);
template <class T>
hipError_t enqueue_control(const Args<T> &A, int observe, hipStream_t st);
template <class T>
long long adj_blocks(const Args<T> &A);
template <class T>
hipError_t enqueue_adj_stage(const Args<T> &A, const AdjArgs<T> &J, int j, hipStream_t st);
template <class T>
hipError_t enqueue_adj_init(const Args<T> &A, const AdjArgs<T> &J, int nstage, hipStream_t st);
hipError_t enqueue_adj_step(long long *step, long long value, int add, hipStream_t st);
template <class T>
hipError_t enqueue_adj_objective(const Args<T> &A, const AdjArgs<T> &J, hipStream_t st);
template <class T>
hipError_t adj_configure(const Args<T> &A);
hipError_t enqueue_adj_stage_f64(const Args<float> &A, const Args<double> &F, const AdjArgs<float> &J, int j,
                                 hipStream_t st);
hipError_t adj_configure_f64(const Args<double> &F);
template <class T>
hipError_t enqueue_kind(const Args<T> &A, int kind, int s, hipStream_t st);
template <class T>
hipError_t enqueue_preview(const T *field, long long ncell, const PreviewGeom &g, void *out, int out_fp64,
                           hipStream_t st);
hipError_t preview_scratch(int device, void **ptr);
bool yfast_factors(int py, int *r1, int *r2);
bool zfast_factors(int pz, int *r1, int *r2);
std::vector<double> newell_reduced_spectrum(int nx, int ny, int nz, double dx, double dy, double dz);
bool newell_periodic_spectrum(int nx, int ny, int nz, double dx, double dy, double dz, const int per[3],
                              const int img[3], std::vector<double> &out, std::string *err);
std::vector<double> newell_layer_spectrum(int nx, int ny, double dx, double dy, double dz, double s);
bool reduce_given_spectra(const double *const spec[6], uint64_t len, int nx, int ny, int nz, std::vector<double> &out,
                          std::string *err);
}  // namespace fm

namespace {
constexpr double kMu0 = 4.0 * 3.14159265358979323846 * 1e-7;
constexpr int kAttemptChunk = 8;  // trial steps enqueued per completion check in advance()
}  // namespace

struct fullmag_fdm_backend {
    std::string err;
    bool ok = false;
    int device = 0;
    hipStream_t stream = nullptr;
    // readout stream (non-blocking, default priority): device-to-host copies
    // of snapshots run here, so the compute stream's next step does not
    // queue behind them
    hipStream_t io_stream = nullptr;
    void *preview_d = nullptr;  // preview bins of copy_field_preview_* (grown on demand)
    size_t preview_cap = 0;
    bool fp64 = false;
    uint32_t nx = 0, ny = 0, nz = 0;
    uint64_t n = 0;
    int nstage = 0;
    bool adaptive = false;
    bool multistep = false;  // ABM3: RHS history in k_all slots kHistSlot.. (Ctl::abm_*)
    bool thermal = false;    // thermal field: RHS depends on Ctl::step and the attempt's dt
    bool fsal = false;       // DP45: k1 is the last stage of the previous step
    fm::Args<float> af{};
    fm::Args<double> ad{};
    std::vector<void *> allocs;
    fm::Ctl *ctl_d = nullptr;
    fm::Ctl *ctl_h = nullptr;  // pinned mirror
    std::vector<uint8_t> mask_h;
    bool has_mask = false;
    double hext[3] = {0, 0, 0};
    bool has_ext = false;
    std::vector<uint8_t> wave_mask_h;  // cells of the Zeeman waveform (empty: all magnetic cells)
    fullmag_fdm_interrupt_poll_fn poll = nullptr;
    void *poll_user = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraphExec_t graph_chunk = nullptr;  // kAttemptChunk trial steps in one graph
    bool graph_tried = false;
    size_t lds_limit = 0;                  // dynamic LDS available to one work-group
    bool obs_fresh = false;
    void *obs_base[3] = {nullptr, nullptr, nullptr};  // hex, hd, heff
    // device copies of the kernel arguments: 0 stepping, 1 with observable
    // outputs, 2 with trajectory recording (adjoint)
    void *args_d[3] = {nullptr, nullptr, nullptr};
    // adjoint sweep (fullmag_fdm_adjoint_*)
    struct Adj {
        std::vector<void *> bufs;
        void *cp = nullptr;                       // ncp x 3N checkpoints
        void *rec_y = nullptr, *rec_h = nullptr, *rec_vn = nullptr;
        void *cot_m = nullptr, *cot_k = nullptr, *beta = nullptr, *obj_w = nullptr, *m_end = nullptr;
        double *grad_ms = nullptr, *grad_amp = nullptr, *obj_part = nullptr, *tn = nullptr;
        long long *step = nullptr;
        uint64_t steps = 0, seg = 0, ncp = 0, first = 0;
        uint64_t krec = 1;      // trailing segments the forward records
        // record ring: rec_seg = krec*seg + 1 step slots (step n in slot n mod
        // rec_seg; a step's end state starts the next slot), so any window of
        // krec segments and its end state, or one segment and its end state,
        // has a slot per step
        uint64_t rec_seg = 0;
        uint32_t comp = 0;
        double dt = 0.0, t0 = 0.0, obj_scale = 0.0;
        long long blocks = 0;
        int parity0 = 0;
        fm::Ctl ctl_end{};
        bool ready = false;
        bool beta64 = false;     // FP32 system: field cotangent convolved in the FP64 context f64
        bool last_rec = false;   // the forward's record of the last krec segments is intact
    } adj;
    // FP64 reduced Newell spectrum of an FP32 demag plan (kept up to 64 MB),
    // and the FP64 demag context built from it on the first adjoint run
    std::vector<double> red64;
    fm::Args<double> f64{};
    bool f64_ready = false;
};

// Asynchronous readouts (reference AsyncFieldSnapshot/AsyncPreviewSnapshot,
// native/backends/fdm/src/context.cu:1271-1745): the payload is staged on
// the device on the compute stream (a device-to-device copy of the field, or
// the preview bins), `ready` is recorded there, and the device-to-host copy
// into pinned memory runs on the backend's readout stream after `ready`,
// ending with `done`.  The snapshot owns its staging buffer, pinned payload
// and events, so it stays valid after the backend is destroyed.
struct fullmag_fdm_field_snapshot {
    void *host = nullptr;     // pinned payload
    void *staging = nullptr;  // device copy of the field
    uint64_t bytes = 0;
    fullmag_fdm_snapshot_desc desc{};
    hipEvent_t ready = nullptr, done = nullptr;
    bool waited = false;
};

struct fullmag_fdm_preview_snapshot {
    void *host = nullptr;     // pinned payload
    void *staging = nullptr;  // device preview bins
    uint64_t bytes = 0;
    fullmag_fdm_snapshot_desc desc{};
    hipEvent_t ready = nullptr, done = nullptr;
    bool waited = false;
};

namespace {

using Backend = fullmag_fdm_backend;

template <class F>
auto with_args(Backend *h, F &&f) {
    return h->fp64 ? f(h->ad) : f(h->af);
}

bool hip_ok(Backend *h, hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    h->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

// step + count without wrap-around (a caller may pass UINT64_MAX as "no limit")
uint64_t sat_add(uint64_t step, uint64_t count) {
    return count > std::numeric_limits<uint64_t>::max() - step ? std::numeric_limits<uint64_t>::max() : step + count;
}

int fail(Backend *h, int code, const std::string &msg) {
    h->err = msg;
    return code;
}

void *dalloc(Backend *h, size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    h->allocs.push_back(p);
    return p;
}

// Reduced Newell spectrum of a grid, computed once per process for each of
// the last few geometries (batched sweeps and repeated creates share it).
std::vector<double> cached_newell(int nx, int ny, int nz, double dx, double dy, double dz) {
    static std::mutex mu;
    static std::vector<std::pair<std::vector<double>, std::shared_ptr<const std::vector<double>>>> cache;
    const std::vector<double> key = {(double)nx, (double)ny, (double)nz, dx, dy, dz};
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto &e : cache)
            if (e.first == key) return *e.second;
    }
    auto v = std::make_shared<const std::vector<double>>(fm::newell_reduced_spectrum(nx, ny, nz, dx, dy, dz));
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() >= 8) cache.erase(cache.begin());
    cache.emplace_back(key, v);
    return *v;
}

// Process-wide x-tile choices of the tuner (fullmag_fdm_backend_autotune),
// keyed by precision, grid and method; applied by later creates.
struct TuneEntry {
    std::vector<int> key;
    int rows, xfast;
    int yl = 0, zf = 0;  // 3-D y-line threads / z-pass ky per work-group (0: default)
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tune;

// Kernel instantiation of a plan (further terms, thermal field, region
// table, demag): plans that differ in it keep their own tuned tile, so that
// a plan created alone uses the tile its batch group was tuned to.
template <class T>
int tune_feat(const fm::Args<T> &A) {
    return (A.xt ? 1 : 0) | (A.has_th ? 2 : 0) | (A.region ? 4 : 0) | (A.demag ? 8 : 0) | (A.lmix ? 16 : 0) |
           (A.layered ? 32 : 0) | (A.pbc << 6) | (A.gtens ? 512 : 0);
}

void tuned_lookup(bool fp64, int nx, int ny, int nz, int method, int feat, int *rows, int *xfast, int *yl = nullptr,
                  int *zf = nullptr) {
    const std::vector<int> key = {fp64 ? 1 : 0, nx, ny, nz, method, feat};
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &e : g_tune)
        if (e.key == key) {
            *rows = e.rows;
            *xfast = e.xfast;
            if (yl) *yl = e.yl;
            if (zf) *zf = e.zf;
        }
}

void tuned_store(bool fp64, int nx, int ny, int nz, int method, int feat, int rows, int xfast, int yl = 0,
                 int zf = 0) {
    const std::vector<int> key = {fp64 ? 1 : 0, nx, ny, nz, method, feat};
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto &e : g_tune)
        if (e.key == key) {
            e.rows = rows;
            e.xfast = xfast;
            e.yl = yl;
            e.zf = zf;
            return;
        }
    g_tune.push_back({key, rows, xfast, yl, zf});
}

// Publishes B as the kernel arguments in device slot `slot` (the kernels
// read their arguments from device memory).  The stream is drained first so
// that no queued launch still reads the slot.
template <class T>
bool put_args(Backend *h, fm::Args<T> &B, int slot) {
    B.self_d = static_cast<const fm::Args<T> *>(h->args_d[slot]);
    B.nsys = 1;
    return hip_ok(h, hipStreamSynchronize(h->stream), "stream sync") &&
           hip_ok(h, hipMemcpy(h->args_d[slot], &B, sizeof(B), hipMemcpyHostToDevice), "kernel arguments");
}

bool plan_fft(int n, fm::FftDesc &fd, std::string *err) {
    fd.n = n;
    fd.npass = 0;
    int m = n;
    auto push = [&](int r) {
        if (fd.npass >= fm::kMaxPasses) return false;
        fd.radix[fd.npass++] = r;
        return true;
    };
    while (m % 8 == 0) {
        if (!push(8)) return false;
        m /= 8;
    }
    while (m % 4 == 0) {
        if (!push(4)) return false;
        m /= 4;
    }
    while (m % 2 == 0) {
        if (!push(2)) return false;
        m /= 2;
    }
    fd.lg = -1;
    if (m == 1) {
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        fd.lg = lg;
    }
    for (int p = 3; m > 1; p += 2) {
        while (m % p == 0) {
            if (p > fm::kMaxRadix) {
                *err = "padded FFT length " + std::to_string(n) + " has a prime factor " + std::to_string(p) +
                       " > " + std::to_string(fm::kMaxRadix);
                return false;
            }
            if (!push(p)) return false;
            m /= p;
        }
    }
    return true;
}

template <class T>
void *upload_twiddles(Backend *h, int n) {
    std::vector<T> tw(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const double a = 2.0 * 3.14159265358979323846 * (double)k / (double)n;
        tw[2 * k] = (T)std::cos(a);
        tw[2 * k + 1] = (T)(-std::sin(a));
    }
    void *d = dalloc(h, tw.size() * sizeof(T));
    if (d && hipMemcpy(d, tw.data(), tw.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

// Four-step twiddle table W_n^{t1 k2} at [k2*r1 + t1] of a register
// four-step length n (none, and true, when n has no such factorisation);
// swap: the table of the transposed factorisation n = r2*r1 (the forward
// 3-D y-line pass runs with the factors exchanged).
template <class T>
bool four_step_twiddles(Backend *h, int n, const void **dst, bool zshape = false, bool swap = false) {
    int r1 = 0, r2 = 0;
    if (!(zshape ? fm::zfast_factors(n, &r1, &r2) : fm::yfast_factors(n, &r1, &r2))) return true;
    if (swap) std::swap(r1, r2);
    std::vector<T> t4(2 * (size_t)n);
    for (int k2 = 0; k2 < r2; ++k2)
        for (int t1 = 0; t1 < r1; ++t1) {
            const double a = 2.0 * 3.14159265358979323846 * (double)((long)t1 * k2) / (double)n;
            t4[2 * ((size_t)k2 * r1 + t1)] = (T)std::cos(a);
            t4[2 * ((size_t)k2 * r1 + t1) + 1] = (T)(-std::sin(a));
        }
    void *d = dalloc(h, t4.size() * sizeof(T));
    if (!d || hipMemcpy(d, t4.data(), t4.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return hip_ok(h, hipErrorOutOfMemory, "four-step twiddles");
    *dst = d;
    return true;
}

// Per-lane twiddle bases of the three-step length-1024 FFTs (the FP32
// film's y pass k_yf8 and x passes at px = 1024): [q][lane] (q < 6,
// lane < 128) holds W_1024^{x} (complex<T>, W =
// exp(-2 pi i / 1024)) with x = 8 t1' e, 16 t1' (steps A), t1 k2', 16 t1
// (B), k2 (C'), 8 k2' (B') for the lane's roles e = lane & 1, t1' = lane >> 4
// in steps A/A', t1 = lane & 7, k2' = lane >> 3 in steps B/B', k2 = lane in
// steps C/C'; the kernel forms the other twiddles as powers of these.  A
// wave's loads of one q are contiguous.
template <class T>
bool f8_bases(Backend *h, const void **dst) {
    const int n = 1024;
    std::vector<T> t(2 * 6 * 128);
    for (int q = 0; q < 6; ++q)
        for (int l = 0; l < 128; ++l) {
            const int e = l & 1, at1p = l >> 4, bt1 = l & 7, bk2p = l >> 3, k2 = l;
            const int xs[6] = {8 * at1p * e, 16 * at1p, bt1 * bk2p, 16 * bt1, k2, 8 * bk2p};
            const double a = 2.0 * 3.14159265358979323846 * (double)(xs[q] % n) / (double)n;
            const size_t i = 2 * ((size_t)q * 128 + l);
            t[i] = (T)std::cos(a);
            t[i + 1] = (T)(-std::sin(a));
        }
    void *d = dalloc(h, t.size() * sizeof(T));
    if (!d || hipMemcpy(d, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return hip_ok(h, hipErrorOutOfMemory, "three-step FFT twiddles");
    *dst = d;
    return true;
}

int select_device(Backend *h) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(h, FULLMAG_FDM_ERR_CUDA, "no HIP device");
    int dev = 0;
    const char *e = std::getenv("FULLMAG_FDM_GPU_INDEX");
    if (!e) e = std::getenv("FULLMAG_CUDA_DEVICE_INDEX");
    if (e) dev = std::atoi(e);
    if (dev < 0 || dev >= count) return fail(h, FULLMAG_FDM_ERR_INVALID, "GPU index out of range");
    h->device = dev;
    if (hipSetDevice(dev) != hipSuccess) return fail(h, FULLMAG_FDM_ERR_CUDA, "hipSetDevice failed");
    return FULLMAG_FDM_OK;
}

// Boundary-correction tiers (fullmag_fdm.h:222-258; uploaded as
// native/backends/fdm/src/api.cpp:436-476 and context.cu:722-800).  The
// volume fraction phi weights the demag source (phi Ms m) and every cell's
// energy; T0 (VOLUME with all six face-link arrays) and T1 (FULL with all six
// boundary distances) replace the exchange stencil by per-face weights
// computed here: T0 f / max(phi, floor) for a face with link f > 0 inside the
// box, T1 2 d / (d + max(delta_opposite, delta_min)) / max(phi, floor) for an
// active neighbour inside the box (exchange_t0_fp64.cu:52-165,
// exchange_t1_fp64.cu:62-198; z faces only for nz > 1; cells with phi <= 0
// get none).  Without the stencil arrays the standard stencil stays, as in
// the reference.  The sparse demag correction becomes a per-target list
// (repeated targets add up; inactive targets and negative indices are
// skipped).
template <class T>
bool parse_boundary(Backend *h, fm::Args<T> &A, const fullmag_fdm_plan_desc *p, const fullmag_fdm_plan_ext *x,
                    double ms) {
    const uint64_t n = h->n;
    A.ms_pack = A.ms_cell;
    const int tier = (int)p->boundary_correction;
    if (tier == FULLMAG_FDM_BOUNDARY_NONE) {
        if (p->has_demag_boundary_corr && p->demag_corr_target_count > 0)
            return (h->err = "sparse demag boundary correction needs a boundary_correction tier with volume_fraction",
                    false);
        return true;
    }
    if (tier != FULLMAG_FDM_BOUNDARY_VOLUME && tier != FULLMAG_FDM_BOUNDARY_FULL)
        return (h->err = "boundary_correction must be 0 (none), 1 (volume) or 2 (full)", false);
    if (!p->volume_fraction || p->volume_fraction_len != n)
        return (h->err = "boundary_correction needs volume_fraction with cell_count values", false);
    for (uint64_t i = 0; i < n; ++i)
        if (!(p->volume_fraction[i] >= 0.0 && p->volume_fraction[i] <= 1.0))
            return (h->err = "volume_fraction must lie in [0, 1]", false);
    const double floor_ = p->boundary_phi_floor > 0.0 ? p->boundary_phi_floor : 0.05;
    const double dmin = p->boundary_delta_min > 0.0 ? p->boundary_delta_min
                                                    : 0.1 * std::min({p->grid.dx, p->grid.dy, p->grid.dz});
    auto upload = [&](const void *src, size_t bytes, const char *what) -> void * {
        void *d = dalloc(h, bytes);
        if (!d || hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            h->err = std::string("upload ") + what;
            return nullptr;
        }
        return d;
    };
    const double *phi = p->volume_fraction;
    std::vector<T> ph(phi, phi + n), mp(n);
    for (uint64_t i = 0; i < n; ++i) {
        const double msi = x && x->ms_field && x->ms_field[i] > 0.0 ? x->ms_field[i] : ms;
        mp[i] = (T)(msi * phi[i]);
    }
    if (!(A.bc_phi = static_cast<const T *>(upload(ph.data(), n * sizeof(T), "volume_fraction"))) ||
        !(A.ms_pack = static_cast<const T *>(upload(mp.data(), n * sizeof(T), "volume_fraction"))))
        return false;
    const double *fl[6] = {p->face_link_xm, p->face_link_xp, p->face_link_ym,
                           p->face_link_yp, p->face_link_zm, p->face_link_zp};
    const double *dl[6] = {p->delta_xm, p->delta_xp, p->delta_ym, p->delta_yp, p->delta_zm, p->delta_zp};
    bool faces = true, dists = true;
    for (int q = 0; q < 6; ++q) {
        faces = faces && fl[q];
        dists = dists && dl[q];
    }
    const int st = tier == FULLMAG_FDM_BOUNDARY_VOLUME && faces ? 1 : tier == FULLMAG_FDM_BOUNDARY_FULL && dists ? 2 : 0;
    if (st) {
        const long nx = h->nx, ny = h->ny, nz = h->nz;
        const double d[3] = {p->grid.dx, p->grid.dy, p->grid.dz};
        const long stride[3] = {1, nx, nx * ny}, ext[3] = {nx, ny, nz};
        std::vector<T> w(6 * n, T(0));
        for (long i = 0; i < (long)n; ++i) {
            if (!(phi[i] > 0.0)) continue;
            const double inv = 1.0 / std::max(phi[i], floor_);
            const long c[3] = {i % nx, (i / nx) % ny, i / (nx * ny)};
            for (int a = 0; a < 3; ++a) {
                if (a == 2 && nz == 1) continue;
                for (int side = 0; side < 2; ++side) {  // 0: -a neighbour, 1: +a neighbour
                    const int q = 2 * a + side;
                    if (side == 0 ? c[a] == 0 : c[a] + 1 >= ext[a]) continue;
                    double v;
                    if (st == 1) {
                        v = fl[q][i] > 0.0 ? fl[q][i] : 0.0;
                    } else {
                        const long j = side == 0 ? i - stride[a] : i + stride[a];
                        if (h->has_mask && !h->mask_h[j]) continue;
                        v = 2.0 * d[a] / (d[a] + std::max(dl[2 * a + 1 - side][i], dmin));
                    }
                    w[(size_t)q * n + i] = (T)(v * inv);
                }
            }
        }
        if (!(A.bc_w = static_cast<const T *>(upload(w.data(), w.size() * sizeof(T), "boundary stencil")))) return false;
    }
    if (p->has_demag_boundary_corr && p->demag_corr_target_count > 0) {
        const uint64_t tc = p->demag_corr_target_count, ss = p->demag_corr_stencil_size;
        if (!p->demag_corr_target_idx || !p->demag_corr_source_idx || !p->demag_corr_tensor || ss == 0)
            return (h->err = "sparse demag boundary correction: target, source and tensor arrays are required", false);
        std::vector<int> cnt(n + 1, 0);
        auto used = [&](uint64_t t) {
            const int32_t ti = p->demag_corr_target_idx[t];
            return ti >= 0 && (!h->has_mask || h->mask_h[ti]);
        };
        for (uint64_t t = 0; t < tc; ++t) {
            const int32_t ti = p->demag_corr_target_idx[t];
            if (ti >= (int64_t)n) return (h->err = "sparse demag boundary correction: target index out of range", false);
            for (uint64_t j = 0; j < ss; ++j)
                if (p->demag_corr_source_idx[t * ss + j] >= (int64_t)n)
                    return (h->err = "sparse demag boundary correction: source index out of range", false);
            if (!used(t)) continue;
            for (uint64_t j = 0; j < ss; ++j) cnt[ti + 1] += p->demag_corr_source_idx[t * ss + j] >= 0;
        }
        for (uint64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
        std::vector<int> src(std::max(cnt[n], 1)), fill(cnt.begin(), cnt.end() - 1);
        std::vector<T> ten(6 * (size_t)std::max(cnt[n], 1));
        for (uint64_t t = 0; t < tc; ++t) {
            if (!used(t)) continue;
            const int32_t ti = p->demag_corr_target_idx[t];
            for (uint64_t j = 0; j < ss; ++j) {
                const int32_t sj = p->demag_corr_source_idx[t * ss + j];
                if (sj < 0) continue;
                const int k = fill[ti]++;
                src[k] = sj;
                for (int c = 0; c < 6; ++c) ten[6 * (size_t)k + c] = (T)p->demag_corr_tensor[6 * (t * ss + j) + c];
            }
        }
        if (!(A.dc_ptr = static_cast<const int *>(upload(cnt.data(), cnt.size() * sizeof(int), "demag correction"))) ||
            !(A.dc_src = static_cast<const int *>(upload(src.data(), src.size() * sizeof(int), "demag correction"))) ||
            !(A.dc_n = static_cast<const T *>(upload(ten.data(), ten.size() * sizeof(T), "demag correction"))))
            return false;
    }
    A.xt = 1;
    return true;
}

// Cubic anisotropy, DMI, magnetoelastic coupling, Zhang-Li / Slonczewski /
// spin-orbit torques and the Oersted cylinder: activation and prefactors as
// native/backends/fdm/src/api.cpp:220-318 (physical constants as there);
// prefactors exclude 1/Ms, applied per cell by the kernels.
template <class T>
bool parse_further_terms(Backend *h, fm::Args<T> &A, const fullmag_fdm_plan_desc *p) {
    constexpr double kMuB = 9.274009994e-24, kE = 1.60217662e-19, kHbar = 1.054571817e-34;
    const auto unit = [](const double *a, double *o) {
        const double n = std::max(std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]), 1e-30);
        for (int c = 0; c < 3; ++c) o[c] = a[c] / n;
    };
    const uint64_t n = h->n;
    if (p->has_cubic_anisotropy) {
        double c1[3], c2[3];
        unit(p->cubic_axis1, c1);
        unit(p->cubic_axis2, c2);
        const double c3[3] = {c1[1] * c2[2] - c1[2] * c2[1], c1[2] * c2[0] - c1[0] * c2[2], c1[0] * c2[1] - c1[1] * c2[0]};
        for (int c = 0; c < 3; ++c) {
            A.cub_ax[c] = (T)c1[c];
            A.cub_ax[3 + c] = (T)c2[c];
            A.cub_ax[6 + c] = (T)c3[c];
        }
        A.kc1_u = (T)p->cubic_Kc1;
        A.kc2_u = (T)p->cubic_Kc2;
        A.kc3_u = (T)p->cubic_Kc3;
        A.has_cub = 1;
        const double *src[3] = {p->kc1_field, p->kc2_field, p->kc3_field};
        const T **dst[3] = {&A.kc1_cell, &A.kc2_cell, &A.kc3_cell};
        for (int q = 0; q < 3; ++q) {
            if (!src[q]) continue;
            std::vector<T> v(src[q], src[q] + n);
            T *d = static_cast<T *>(dalloc(h, n * sizeof(T)));
            if (!d || hipMemcpy(d, v.data(), n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
                return (h->err = "upload cubic constant field", false);
            *dst[q] = d;
        }
    }
    if (p->has_interfacial_dmi) A.dmi_i = (T)p->dmi_D_interfacial;
    if (p->has_bulk_dmi) A.dmi_b = (T)p->dmi_D_bulk;
    A.i2x = (T)(0.5 / p->grid.dx);
    A.i2y = (T)(0.5 / p->grid.dy);
    A.i2z = (T)(0.5 / p->grid.dz);
    A.i1x = (T)(1.0 / p->grid.dx);
    A.i1y = (T)(1.0 / p->grid.dy);
    A.i1z = (T)(1.0 / p->grid.dz);
    if (p->has_magnetoelastic) {
        // Voigt [e11 e22 e33 2e23 2e13 2e12] -> tensor components
        const double *e = p->mel_strain;
        const double t[6] = {e[0], e[1], e[2], 0.5 * e[3], 0.5 * e[4], 0.5 * e[5]};
        for (int i = 0; i < 6; ++i) A.mel_e[i] = (T)t[i];
        A.mel_b1 = (T)p->mel_b1;
        A.mel_b2 = (T)p->mel_b2;
        A.has_mel = 1;
    }
    const double j[3] = {p->current_density_x, p->current_density_y, p->current_density_z};
    const double jm = std::sqrt(j[0] * j[0] + j[1] * j[1] + j[2] * j[2]);
    if (jm > 0.0 && p->stt_degree > 0.0) {
        const double beta = p->stt_beta;
        const double b = p->stt_degree * kMuB / (kE * (1.0 + beta * beta));
        for (int c = 0; c < 3; ++c) A.zl_u[c] = (T)(b * j[c]);
        A.zl_beta = (T)beta;
        A.has_zl = 1;
    }
    const double pp = p->stt_p_x * p->stt_p_x + p->stt_p_y * p->stt_p_y + p->stt_p_z * p->stt_p_z;
    if (pp > 0.0 && p->stt_lambda > 0.0 && jm > 0.0) {
        A.sl_p[0] = (T)p->stt_p_x;
        A.sl_p[1] = (T)p->stt_p_y;
        A.sl_p[2] = (T)p->stt_p_z;
        A.sl_l2 = (T)(p->stt_lambda * p->stt_lambda);
        A.sl_deg = (T)(p->stt_degree > 0.0 ? p->stt_degree : 1.0);
        A.sl_eps = (T)p->stt_epsilon_prime;
        A.sl_pf = (T)(jm * kHbar / (2.0 * kE * kMu0 * p->grid.dz));
        A.has_slon = 1;
    }
    if (p->has_sot) {
        if (!(p->sot_thickness > 0.0)) return (h->err = "sot_thickness must be positive", false);
        if (p->sot_je != 0.0) {
            double sg[3];
            unit(p->sot_sigma, sg);
            for (int c = 0; c < 3; ++c) A.sot_s[c] = (T)sg[c];
            A.sot_pf = (T)(std::fabs(p->sot_je) * kHbar / (2.0 * kE * kMu0 * p->sot_thickness));
            A.sot_dl = (T)p->sot_xi_dl;
            A.sot_fl = (T)p->sot_xi_fl;
            A.has_sot = 1;
        }
    }
    if (p->has_oersted_cylinder) {
        if (!(p->oersted_radius > 0.0)) return (h->err = "oersted_radius must be positive", false);
        if (p->oersted_time_dep_kind > 2) return (h->err = "oersted_time_dep_kind must be 0, 1 or 2", false);
        // profile at the full current (fdm_fields.rs:1049-1150 geometry:
        // cell centres, distance from the axis line, azimuthal axis x r_hat)
        double ax[3];
        unit(p->oersted_axis, ax);
        const double I = p->oersted_current, R = p->oersted_radius;
        std::vector<T> prof(3 * n, (T)0);
        const uint64_t nx = h->nx, ny = h->ny;
        for (uint64_t i = 0; i < n; ++i) {  // (empty cells never read it)
            const uint64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
            const double r0[3] = {(x + 0.5) * p->grid.dx - p->oersted_center[0],
                                  (y + 0.5) * p->grid.dy - p->oersted_center[1],
                                  (z + 0.5) * p->grid.dz - p->oersted_center[2]};
            const double pr = r0[0] * ax[0] + r0[1] * ax[1] + r0[2] * ax[2];
            const double r[3] = {r0[0] - pr * ax[0], r0[1] - pr * ax[1], r0[2] - pr * ax[2]};
            const double rp = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (rp < 1e-30) continue;
            const double hm = rp <= R ? I * rp / (2.0 * M_PI * R * R) : I / (2.0 * M_PI * rp);
            const double phi[3] = {(ax[1] * r[2] - ax[2] * r[1]) / rp, (ax[2] * r[0] - ax[0] * r[2]) / rp,
                                   (ax[0] * r[1] - ax[1] * r[0]) / rp};
            for (int c = 0; c < 3; ++c) prof[c * n + i] = (T)(phi[c] * hm);
        }
        T *d = static_cast<T *>(dalloc(h, 3 * n * sizeof(T)));
        if (!d || hipMemcpy(d, prof.data(), prof.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return (h->err = "upload Oersted field", false);
        A.oe_field = d;
        A.oe_kind = (int)p->oersted_time_dep_kind;
        A.oe_freq = p->oersted_time_dep_freq;
        A.oe_phase = p->oersted_time_dep_phase;
        A.oe_offset = p->oersted_time_dep_offset;
        A.oe_t_on = p->oersted_time_dep_t_on;
        A.oe_t_off = p->oersted_time_dep_t_off;
    }
    A.xt = (A.has_cub || A.dmi_i != (T)0 || A.dmi_b != (T)0 || A.has_mel || A.has_zl || A.has_slon || A.has_sot ||
            A.oe_field)
               ? 1
               : 0;
    return true;
}

template <class T>
bool upload_state(Backend *h, const double *aos_d, const float *aos_f, int slot) {
    auto &A = *reinterpret_cast<fm::Args<T> *>(h->fp64 ? (void *)&h->ad : (void *)&h->af);
    const uint64_t n = h->n;
    std::vector<T> soa(3 * n);
    for (uint64_t i = 0; i < n; ++i) {
        double v[3];
        for (int c = 0; c < 3; ++c) v[c] = aos_d ? aos_d[3 * i + c] : (double)aos_f[3 * i + c];
        const bool act = !h->has_mask || h->mask_h[i] != 0;
        const double nrm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int c = 0; c < 3; ++c) soa[c * n + i] = (act && nrm > 0.0) ? (T)(v[c] * (1.0 / nrm)) : (T)0;
    }
    return hip_ok(h, hipMemcpy(A.m_all + (size_t)slot * 3 * n, soa.data(), soa.size() * sizeof(T), hipMemcpyHostToDevice),
                  "upload magnetisation");
}

bool push_ctl(Backend *h) {
    return hip_ok(h, hipMemcpyAsync(h->ctl_d, h->ctl_h, sizeof(fm::Ctl), hipMemcpyHostToDevice, h->stream), "ctl upload");
}

bool pull_ctl(Backend *h) {
    if (!hip_ok(h, hipMemcpyAsync(h->ctl_h, h->ctl_d, sizeof(fm::Ctl), hipMemcpyDeviceToHost, h->stream), "ctl download"))
        return false;
    return hip_ok(h, hipStreamSynchronize(h->stream), "stream sync");
}

// k1 of the current state and its statistics (stage-1 evaluation).
// With a thermal field the evaluation takes the noise of the current counter
// at the dt of the last accepted step (the CPU engine's fields between
// steps), and the FSAL stage of DP45, whose noise belongs to the previous
// counter, is kept: it is copied to the stage-2 slot, which is free between
// steps, and back.
bool eval_current(Backend *h, bool observables) {
    double dt_keep = 0.0;
    bool keep_k1 = false;
    if (h->thermal) {
        if (!pull_ctl(h)) return false;
        dt_keep = h->ctl_h->dt;
        h->ctl_h->dt = h->ctl_h->dt_used;
        if (!push_ctl(h)) return false;
        keep_k1 = h->fsal && !h->ctl_h->k1_redo;
    }
    auto k1_copy = [&](bool save) {
        return with_args(h, [&](auto &A) {
            const size_t v = 3 * (size_t)h->n * sizeof(*A.k_all);
            auto *k1 = A.k_all + (size_t)h->ctl_h->parity * 3 * h->n, *k2 = A.k_all + 2 * 3 * h->n;
            return hip_ok(h, hipMemcpyAsync(save ? k2 : k1, save ? k1 : k2, v, hipMemcpyDeviceToDevice, h->stream),
                          "keep k1");
        });
    };
    if (keep_k1 && !k1_copy(true)) return false;
    const bool ok = with_args(h, [&](auto &A) {
        auto B = A;
        if (observables) {
            using T = typename std::remove_pointer<decltype(A.m_all)>::type;
            B.obs_hex = reinterpret_cast<T *>(h->obs_base[0]);
            B.obs_hd = reinterpret_cast<T *>(h->obs_base[1]);
            B.obs_heff = reinterpret_cast<T *>(h->obs_base[2]);
            if (!put_args(h, B, 1)) return false;
        }
        return hip_ok(h, fm::enqueue_eval(B, 1, 2, h->stream), "launch (evaluation)");
    });
    if (!ok || (keep_k1 && !k1_copy(false)) || !pull_ctl(h)) return false;
    if (h->thermal) {
        h->ctl_h->dt = dt_keep;
        if (!push_ctl(h)) return false;
    }
    if (observables) h->obs_fresh = true;
    return true;
}

bool enqueue_attempt(Backend *h) {
    return with_args(h, [&](auto &A) {
        return hip_ok(h, fm::enqueue_trial(A, h->nstage, h->stream), "launch (trial step)");
    });
}

// Captures one trial step as a HIP graph (falls back to direct launches).
void ensure_graph(Backend *h) {
    if (h->graph_tried) return;
    h->graph_tried = true;
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return;
    const bool ok = enqueue_attempt(h);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (!ok || e != hipSuccess || !g) {
        if (g) hipGraphDestroy(g);
        (void)hipGetLastError();
        h->err.clear();
        return;
    }
    hipGraphExec_t ex = nullptr;
    if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
        hipGraphDestroy(g);
        (void)hipGetLastError();
        return;
    }
    h->graph = g;
    h->graph_exec = ex;
    // the device loop's chunks as one graph: the kernels of attempts after
    // the loop finished return at once
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return;
    bool ok2 = true;
    for (int a = 0; a < kAttemptChunk && ok2; ++a) ok2 = enqueue_attempt(h);
    hipGraph_t g2 = nullptr;
    const hipError_t e2 = hipStreamEndCapture(h->stream, &g2);
    if (ok2 && e2 == hipSuccess && g2 && hipGraphInstantiate(&h->graph_chunk, g2, nullptr, nullptr, 0) != hipSuccess)
        h->graph_chunk = nullptr;
    if (g2) hipGraphDestroy(g2);
    (void)hipGetLastError();
    h->err.clear();
}

bool launch_attempt(Backend *h) {
    if (h->graph_exec) return hip_ok(h, hipGraphLaunch(h->graph_exec, h->stream), "graph launch");
    return enqueue_attempt(h);
}

void fill_stats(Backend *h, fullmag_fdm_step_stats *s, double dt, uint64_t wall_ns) {
    if (!s) return;
    const fm::Ctl &c = *h->ctl_h;
    std::memset(s, 0, sizeof(*s));
    s->step = c.step;
    s->time_seconds = c.t;
    s->dt_seconds = dt;
    s->exchange_energy_joules = c.stats[fm::ST_EEX];
    s->demag_energy_joules = c.stats[fm::ST_EDEMAG];
    s->external_energy_joules = c.stats[fm::ST_EEXT];
    s->anisotropy_energy_joules = c.stats[fm::ST_EANI];
    s->cubic_energy_joules = c.stats[fm::ST_ECUB];
    s->dmi_energy_joules = c.stats[fm::ST_EDMI];
    // every evaluated term, magnetoelastic included (it has no field of its own here)
    s->total_energy_joules = c.stats[fm::ST_EEX] + c.stats[fm::ST_EDEMAG] + c.stats[fm::ST_EEXT] +
                             c.stats[fm::ST_EANI] + c.stats[fm::ST_ECUB] + c.stats[fm::ST_EDMI] +
                             c.stats[fm::ST_EMEL];
    s->max_effective_field_amplitude = std::sqrt(c.stats[fm::ST_HEFF2]);
    s->max_demag_field_amplitude = std::sqrt(c.stats[fm::ST_HD2]);
    s->max_rhs_amplitude = std::sqrt(c.stats[fm::ST_RHS2]);
    s->suggested_next_dt = h->adaptive ? c.dt_next : 0.0;
    s->wall_time_ns = wall_ns;
}

template <class T>
bool create_typed(Backend *h, const fullmag_fdm_plan_desc *p, const fullmag_fdm_plan_ext *x) {
    fm::Args<T> &A = *reinterpret_cast<fm::Args<T> *>(h->fp64 ? (void *)&h->ad : (void *)&h->af);
    std::memset(&A, 0, sizeof(A));
    const int nx = (int)h->nx, ny = (int)h->ny, nz = (int)h->nz;
    const uint64_t n = h->n;
    A.nx = nx;
    A.ny = ny;
    A.nz = nz;
    A.ncell = (long long)n;
    A.nrows = (long long)ny * nz;
    A.nkx = nx + 1;
    A.ly = nz == 1 ? ny : 2 * ny;
    A.nzs = nz;
    A.px = 2 * nx;
    A.py = 2 * ny;
    A.pz = nz == 1 ? 1 : 2 * nz;
    A.inv_p = 1.0 / ((double)A.px * A.py * A.pz);
    const double ms = p->material.saturation_magnetisation;
    double a_ex = p->material.exchange_stiffness;
    const double alpha = p->material.damping;
    const double gamma = p->material.gyromagnetic_ratio > 0.0 ? p->material.gyromagnetic_ratio : 2.211e5;
    A.ms = (T)ms;
    A.demag = p->enable_demag ? 1 : 0;
    A.precession = p->disable_precession ? 0 : 1;
    A.gamma_bar = (T)(gamma / (1.0 + alpha * alpha));
    A.alpha = (T)alpha;
    A.idx2 = (T)(1.0 / (p->grid.dx * p->grid.dx));
    A.idy2 = (T)(1.0 / (p->grid.dy * p->grid.dy));
    A.idz2 = (T)(1.0 / (p->grid.dz * p->grid.dz));
    A.e_pre = kMu0 * ms * p->grid.dx * p->grid.dy * p->grid.dz;
    h->has_ext = p->has_external_field != 0;
    A.has_ext = h->has_ext ? 1 : 0;
    for (int c = 0; c < 3; ++c) {
        h->hext[c] = h->has_ext ? p->external_field_am[c] : 0.0;
        A.hext[c] = (T)h->hext[c];
    }
    A.has_uni = p->has_uniaxial_anisotropy ? 1 : 0;
    if (A.has_uni) {
        const double *u = p->anisotropy_axis;
        const double nu = std::max(std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), 1e-30);
        for (int c = 0; c < 3; ++c) A.axis[c] = (T)(u[c] / nu);
        A.uni_c1 = (T)(2.0 * p->uniaxial_anisotropy_constant / (kMu0 * ms));
        A.uni_c2 = (T)(4.0 * p->uniaxial_anisotropy_k2 / (kMu0 * ms));
        A.ku1_u = (T)p->uniaxial_anisotropy_constant;
        A.ku2_u = (T)p->uniaxial_anisotropy_k2;
    }
    A.atol = p->adaptive_max_error > 0 ? p->adaptive_max_error : 1e-5;
    A.dt_min = p->adaptive_dt_min > 0 ? p->adaptive_dt_min : 1e-18;
    A.dt_max = p->adaptive_dt_max > 0 ? p->adaptive_dt_max : 1e-10;
    A.headroom = p->adaptive_headroom > 0 ? p->adaptive_headroom : 0.8;
    A.rtol = 0.0;
    A.growth = std::numeric_limits<double>::infinity();
    A.shrink = 0.0;
    A.mu0_vol = kMu0 * p->grid.dx * p->grid.dy * p->grid.dz;
    A.two_mu0inv = (T)(2.0 / kMu0);
    A.cell_vol = p->grid.dx * p->grid.dy * p->grid.dz;
    if (!parse_further_terms(h, A, p)) return false;
    if (p->temperature > 0.0) {
        // fdm_fields.rs:818-905: sigma^2 = 2 alpha kB T / (gamma (1 + alpha^2) mu0 Ms V dt)
        constexpr double kKb = 1.380649e-23;
        A.has_th = 1;
        A.th_pre = 2.0 * alpha * kKb * p->temperature / (gamma * (1.0 + alpha * alpha) * kMu0 * A.cell_vol);
        const bool v5 = x && x->struct_size >= offsetof(fullmag_fdm_plan_ext, periodic_axes);
        A.th_dt = v5 && x->thermal_dt > 0.0 ? x->thermal_dt : 0.0;
        A.th_seed = v5 && x->thermal_seed ? x->thermal_seed : 42ull;
        A.xt = 1;
    }
    if (x) {
        if (x->adaptive_rtol < 0 || x->adaptive_shrink_limit < 0) {
            h->err = "invalid adaptive tolerances in plan extension";
            return false;
        }
        A.rtol = x->adaptive_rtol;
        A.growth = x->adaptive_growth_limit > 0 ? x->adaptive_growth_limit : std::numeric_limits<double>::infinity();
        A.shrink = x->adaptive_shrink_limit;
        if (x->field_waveform > fm::WAVE_SINE) {
            h->err = "unknown field_waveform";
            return false;
        }
        A.wave_kind = (int)x->field_waveform;
        for (int c = 0; c < 3; ++c) A.wave_amp[c] = x->field_amplitude_am[c];
        A.wave_t0 = x->field_t0;
        A.wave_t1 = x->field_t1;
        A.wave_freq = x->field_frequency;
        A.wave_phase = x->field_phase;
        if (A.wave_kind == fm::WAVE_RAMP && !(x->field_t1 > x->field_t0)) {
            h->err = "ramp waveform needs field_t1 > field_t0";
            return false;
        }
    }
    A.method = p->integrator == FULLMAG_FDM_INTEGRATOR_DP45   ? fm::M_DP45
               : p->integrator == FULLMAG_FDM_INTEGRATOR_RK4  ? fm::M_RK4
               : p->integrator == FULLMAG_FDM_INTEGRATOR_RK23 ? fm::M_RK23
               : p->integrator == FULLMAG_FDM_INTEGRATOR_ABM3 ? fm::M_ABM3
                                                               : fm::M_HEUN;
    h->nstage = fm::method_nstage(A.method);
    for (int s = 0; s <= fm::kMaxStages; ++s) A.stage_nk[s] = fm::method_stage_nk(A.method, s);
    h->adaptive = A.method == fm::M_DP45 || A.method == fm::M_RK23;
    h->multistep = A.method == fm::M_ABM3;
    h->thermal = A.has_th != 0;
    h->fsal = A.method == fm::M_DP45;

    // geometry
    if (p->active_mask) {
        if (p->active_mask_len != n) {
            h->err = "active_mask_len mismatch: expected " + std::to_string(n) + ", got " +
                     std::to_string(p->active_mask_len);
            return false;
        }
        h->has_mask = true;
        h->mask_h.assign(p->active_mask, p->active_mask + n);
        for (auto &v : h->mask_h) v = v ? 1 : 0;
        auto *d = static_cast<uint8_t *>(dalloc(h, n));
        if (!d) return hip_ok(h, hipErrorOutOfMemory, "alloc mask");
        if (!hip_ok(h, hipMemcpy(d, h->mask_h.data(), n, hipMemcpyHostToDevice), "upload mask")) return false;
        A.mask = d;
    }
    // exchange: uniform stiffness unless the region map uses more than one id
    if (p->region_mask) {
        if (p->region_mask_len != n) {
            h->err = "region_mask_len mismatch: expected " + std::to_string(n) + ", got " +
                     std::to_string(p->region_mask_len);
            return false;
        }
        const int R = FULLMAG_FDM_MAX_EXCHANGE_REGIONS;
        if (p->exchange_lut && p->exchange_lut_len != (uint64_t)R * R) {
            h->err = "exchange_lut_len mismatch: expected " + std::to_string(R * R);
            return false;
        }
        std::vector<double> lut((size_t)R * R, 0.0);
        if (p->exchange_lut) std::memcpy(lut.data(), p->exchange_lut, lut.size() * sizeof(double));
        else
            for (int r = 0; r < R; ++r) lut[(size_t)r * R + r] = a_ex;
        bool uniform = true;
        for (uint64_t i = 0; i < n; ++i) {
            if (p->region_mask[i] >= (uint32_t)R) {
                h->err = "region id out of range (>= 256)";
                return false;
            }
            if (p->region_mask[i] != p->region_mask[0]) uniform = false;
        }
        if (uniform) {
            a_ex = lut[(size_t)p->region_mask[0] * R + p->region_mask[0]];
        } else {
            auto *rd = static_cast<uint32_t *>(dalloc(h, n * 4));
            auto *ld = static_cast<T *>(dalloc(h, lut.size() * sizeof(T)));
            if (!rd || !ld) return hip_ok(h, hipErrorOutOfMemory, "alloc regions");
            std::vector<T> lt(lut.begin(), lut.end());
            if (!hip_ok(h, hipMemcpy(rd, p->region_mask, n * 4, hipMemcpyHostToDevice), "upload regions")) return false;
            if (!hip_ok(h, hipMemcpy(ld, lt.data(), lt.size() * sizeof(T), hipMemcpyHostToDevice), "upload lut"))
                return false;
            A.region = rd;
            A.lut = ld;
        }
    }
    A.ex_pre = p->enable_exchange ? (T)(2.0 * a_ex / (kMu0 * ms)) : (T)0;
    A.two_a_mu0 = p->enable_exchange ? (T)(2.0 * a_ex / kMu0) : (T)0;
    // per-cell maps (reference ku1_field/ku2_field; extension Ms and axis)
    auto upload_cell = [&](const double *src, size_t count, const char *what) -> const T * {
        std::vector<T> v(src, src + count);
        T *d = static_cast<T *>(dalloc(h, count * sizeof(T)));
        if (!d || hipMemcpy(d, v.data(), count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
            h->err = std::string("upload ") + what;
            return nullptr;
        }
        return d;
    };
    if (p->ku1_field && !(A.ku1_cell = upload_cell(p->ku1_field, n, "ku1_field"))) return false;
    if (p->ku2_field && !(A.ku2_cell = upload_cell(p->ku2_field, n, "ku2_field"))) return false;
    if (x && x->ms_field) {
        for (uint64_t i = 0; i < n; ++i)
            if ((!h->has_mask || h->mask_h[i]) && !(x->ms_field[i] > 0.0)) {
                h->err = "ms_field must be positive on magnetic cells";
                return false;
            }
        std::vector<double> msv(x->ms_field, x->ms_field + n);
        for (uint64_t i = 0; i < n; ++i)
            if (!(msv[i] > 0.0)) msv[i] = ms;  // empty cells: any positive value (m = 0 there)
        if (!(A.ms_cell = upload_cell(msv.data(), n, "ms_field"))) return false;
    }
    if (x && x->anisotropy_axis_field) {
        std::vector<double> ax(3 * n);
        for (uint64_t i = 0; i < n; ++i) {
            const double *u = x->anisotropy_axis_field + 3 * i;
            const double nu = std::max(std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), 1e-30);
            for (int c = 0; c < 3; ++c) ax[c * n + i] = u[c] / nu;
        }
        if (!(A.axis_cell = upload_cell(ax.data(), 3 * n, "anisotropy_axis_field"))) return false;
    }
    if (x && x->field_mask) {
        std::vector<uint8_t> wm(x->field_mask, x->field_mask + n);
        for (auto &v : wm) v = v ? 1 : 0;
        h->wave_mask_h = wm;
        auto *d = static_cast<uint8_t *>(dalloc(h, n));
        if (!d || hipMemcpy(d, wm.data(), n, hipMemcpyHostToDevice) != hipSuccess) {
            h->err = "upload field_mask";
            return false;
        }
        A.wave_mask = d;
    }
    A.lut_pre = p->enable_exchange ? (T)(2.0 / (kMu0 * ms)) : (T)0;
    // plan-supplied demag spectra (native/backends/fdm/src/api.cpp:207,362-388):
    // a nonzero demag_kernel_spectrum_len selects them, with every component
    // pointer present and the length of the 3-D form, or for nz == 1 of the
    // thin-film 2-D form; checked whether or not demag is enabled
    if (p->demag_kernel_spectrum_len != 0) {
        const double *given[6] = {p->demag_kernel_xx_spectrum, p->demag_kernel_yy_spectrum,
                                  p->demag_kernel_zz_spectrum, p->demag_kernel_xy_spectrum,
                                  p->demag_kernel_xz_spectrum, p->demag_kernel_yz_spectrum};
        for (const double *g : given)
            if (!g)
                return (h->err = "demag kernel spectra pointers must all be present when demag_kernel_spectrum_len "
                                 "is set",
                        false);
        const uint64_t p3 = 2ull * (2ull * nx) * (2ull * ny) * (2ull * nz), p2 = 2ull * (2ull * nx) * (2ull * ny);
        const uint64_t len = p->demag_kernel_spectrum_len;
        if (len != p3 && !(nz == 1 && len == p2))
            return (h->err = "demag_kernel_spectrum_len mismatch: expected " + std::to_string(p3) + " (3D)" +
                             (nz == 1 ? " or " + std::to_string(p2) + " (thin-film 2D)" : std::string()) + ", got " +
                             std::to_string(len),
                    false);
    }
    // periodic boundaries (plan extension version 6; the CPU engine's
    // FdmBoundaryPolicy with demag_image_counts, default 10 per axis)
    int per[3] = {0, 0, 0}, img[3] = {10, 10, 10};
    const bool v6 = x && x->struct_size >= offsetof(fullmag_fdm_plan_ext, layer_z);
    const double *layer_z = x && x->struct_size >= sizeof(fullmag_fdm_plan_ext) ? x->layer_z : nullptr;
    if (v6 && x->periodic_axes) {
        if (x->periodic_axes & ~7u) return (h->err = "periodic_axes: only bits 0 (x), 1 (y), 2 (z)", false);
        if (p->boundary_correction != FULLMAG_FDM_BOUNDARY_NONE)
            return (h->err = "periodic boundaries cannot be combined with boundary-correction tiers", false);
        if (p->demag_kernel_spectrum_len != 0)
            return (h->err = "periodic boundaries need the engine's own demag kernel (plan spectra are open-boundary)",
                    false);
        for (int a = 0; a < 3; ++a) {
            per[a] = (x->periodic_axes >> a) & 1;
            if (x->periodic_images[a]) img[a] = (int)x->periodic_images[a];
            if (per[a] && img[a] > 1000) return (h->err = "periodic_images above 1000", false);
        }
        A.pbc = (int)x->periodic_axes;
        A.xt = 1;
    }
    if (layer_z) {
        if (nz < 2 || nz > fm::kMaxLayers)
            return (h->err = "layer_z needs 2.." + std::to_string(fm::kMaxLayers) + " planes (grid nz)", false);
        for (uint32_t k = 0; k < nz; ++k)
            if (!std::isfinite(layer_z[k])) return (h->err = "layer_z must be finite", false);
        if (per[2]) return (h->err = "layer_z cannot be combined with a periodic z axis", false);
        if (p->boundary_correction != FULLMAG_FDM_BOUNDARY_NONE)
            return (h->err = "layer_z cannot be combined with boundary-correction tiers", false);
        if (p->demag_kernel_spectrum_len != 0)
            return (h->err = "layer_z needs the engine's own demag kernel (plan spectra are a uniform grid)", false);
        A.layered = 1;
        A.xt = 1;
    }
    if (!parse_boundary(h, A, p, x, ms)) return false;
    // activity of each cell and of its six stencil neighbours in one byte,
    // with the kernels' neighbour rule (box clamp, periodic wrap, no z
    // neighbours in a multilayer stack)
    if (h->has_mask) {
        const long nx_ = (long)h->nx, ny_ = (long)h->ny, nz_ = (long)h->nz, sxy = nx_ * ny_;
        std::vector<uint8_t> nm(h->n);
        for (long i = 0; i < (long)h->n; ++i) {
            const long xi = i % nx_, yi = (i / nx_) % ny_, zi = i / sxy;
            long nb[6] = {xi > 0 ? i - 1 : i,        xi < nx_ - 1 ? i + 1 : i,  yi > 0 ? i - nx_ : i,
                          yi < ny_ - 1 ? i + nx_ : i, zi > 0 ? i - sxy : i,       zi < nz_ - 1 ? i + sxy : i};
            if (A.pbc & 1) {
                if (xi == 0) nb[0] = i + nx_ - 1;
                if (xi == nx_ - 1) nb[1] = i - (nx_ - 1);
            }
            if (A.pbc & 2) {
                if (yi == 0) nb[2] = i + (ny_ - 1) * nx_;
                if (yi == ny_ - 1) nb[3] = i - (ny_ - 1) * nx_;
            }
            if (A.pbc & 4) {
                if (zi == 0) nb[4] = i + (nz_ - 1) * sxy;
                if (zi == nz_ - 1) nb[5] = i - (nz_ - 1) * sxy;
            }
            if (A.layered) nb[4] = nb[5] = i;
            uint8_t b = h->mask_h[i] ? 1 : 0;
            for (int q = 0; q < 6; ++q) b |= (uint8_t)((h->mask_h[nb[q]] ? 1 : 0) << (q + 1));
            nm[i] = b;
        }
        auto *d = static_cast<uint8_t *>(dalloc(h, h->n));
        if (!d || !hip_ok(h, hipMemcpy(d, nm.data(), h->n, hipMemcpyHostToDevice), "upload neighbour mask"))
            return d ? false : hip_ok(h, hipErrorOutOfMemory, "alloc neighbour mask");
        A.nmask = d;
    }

    // state buffers
    A.m_all = static_cast<T *>(dalloc(h, 2 * 3 * n * sizeof(T)));
    A.k_all = static_cast<T *>(dalloc(h, 7 * 3 * n * sizeof(T)));
    A.ybuf = static_cast<T *>(dalloc(h, 6 * n * sizeof(T)));  // two stage-state buffers
    if (h->adaptive) {
        // the error estimate reads the final combination's error-term sum
        if (!fm::method_efuse(A.method)) return (h->err = "internal: adaptive tableau without fused error terms", false);
        A.errv = static_cast<double *>(dalloc(h, 3 * n * sizeof(double)));
        if (!A.errv) return hip_ok(h, hipErrorOutOfMemory, "alloc error terms");
    }
    for (int i = 0; i < 3; ++i) h->obs_base[i] = dalloc(h, 3 * n * sizeof(T));
    h->ctl_d = static_cast<fm::Ctl *>(dalloc(h, sizeof(fm::Ctl)));
    if (!A.m_all || !A.k_all || !A.ybuf || !h->obs_base[0] || !h->obs_base[1] || !h->obs_base[2] || !h->ctl_d)
        return hip_ok(h, hipErrorOutOfMemory, "alloc state");
    A.ctl = h->ctl_d;
    if (!hip_ok(h, hipMemset(A.k_all, 0, 7 * 3 * n * sizeof(T)), "clear k")) return false;
    if (!hip_ok(h, hipMemset(A.m_all, 0, 2 * 3 * n * sizeof(T)), "clear m")) return false;

    // demag
    if (A.demag) {
        std::string e;
        if (!plan_fft(A.px, A.fx, &e) || !plan_fft(A.py, A.fy, &e) || !plan_fft(A.pz, A.fz, &e)) {
            h->err = e.empty() ? "FFT plan failed" : e;
            return false;
        }
        std::vector<double> red;
        const double *spec[6] = {p->demag_kernel_xx_spectrum, p->demag_kernel_yy_spectrum, p->demag_kernel_zz_spectrum,
                                 p->demag_kernel_xy_spectrum, p->demag_kernel_xz_spectrum, p->demag_kernel_yz_spectrum};
        if (layer_z) {
            // pair kernels of every (destination, source) plane; the self
            // pairs are the film kernel (shift 0)
            const int L = nz;
            const size_t ps = (size_t)(nx + 1) * (ny + 1);
            std::vector<T> lk((size_t)L * L * 6 * ps);
            for (int l = 0; l < L; ++l)
                for (int q = 0; q < L; ++q) {
                    const std::vector<double> k = fm::newell_layer_spectrum(nx, ny, p->grid.dx, p->grid.dy, p->grid.dz,
                                                                            layer_z[l] - layer_z[q]);
                    for (size_t i = 0; i < 6 * ps; ++i) lk[(size_t)(l * L + q) * 6 * ps + i] = (T)k[i];
                }
            T *ld = static_cast<T *>(dalloc(h, lk.size() * sizeof(T)));
            if (!ld) return hip_ok(h, hipErrorOutOfMemory, "alloc layer kernels");
            if (!hip_ok(h, hipMemcpy(ld, lk.data(), lk.size() * sizeof(T), hipMemcpyHostToDevice), "upload tensor"))
                return false;
            A.lmix = ld;
            A.inv_p = 1.0 / ((double)A.px * A.py);
        } else if (A.pbc) {
            std::vector<double> gs;
            if (!fm::newell_periodic_spectrum(nx, ny, nz, p->grid.dx, p->grid.dy, p->grid.dz, per, img, gs, &e)) {
                h->err = e;
                return false;
            }
            std::vector<T> gt(gs.begin(), gs.end());
            T *gd = static_cast<T *>(dalloc(h, gt.size() * sizeof(T)));
            if (!gd) return hip_ok(h, hipErrorOutOfMemory, "alloc periodic tensor");
            if (!hip_ok(h, hipMemcpy(gd, gt.data(), gt.size() * sizeof(T), hipMemcpyHostToDevice), "upload tensor"))
                return false;
            A.gtens = gd;
        } else if (p->demag_kernel_spectrum_len != 0) {
            if (!fm::reduce_given_spectra(spec, p->demag_kernel_spectrum_len, nx, ny, nz, red, &e)) {
                h->err = e;
                return false;
            }
        } else {
            red = cached_newell(nx, ny, nz, p->grid.dx, p->grid.dy, p->grid.dz);
        }
        std::vector<T> rt(std::max<size_t>(red.size(), 1));
        if (nz == 1) {
            for (size_t i = 0; i < red.size(); ++i) rt[i] = (T)red[i];
        } else if (!red.empty()) {
            // 3-D: [kx][kz][ky][component], the six components of a
            // frequency contiguous for the z pass
            const size_t X = nx + 1, Y = ny + 1, Z = nz + 1;
            for (size_t c = 0; c < 6; ++c)
                for (size_t kx = 0; kx < X; ++kx)
                    for (size_t ky = 0; ky < Y; ++ky)
                        for (size_t kz = 0; kz < Z; ++kz)
                            rt[((kx * Z + kz) * Y + ky) * 6 + c] = (T)red[((c * X + kx) * Y + ky) * Z + kz];
        }
        T *nt = static_cast<T *>(dalloc(h, rt.size() * sizeof(T)));
        if (!nt) return hip_ok(h, hipErrorOutOfMemory, "alloc tensor");
        if (!hip_ok(h, hipMemcpy(nt, rt.data(), rt.size() * sizeof(T), hipMemcpyHostToDevice), "upload tensor"))
            return false;
        A.ntens = nt;
        if (!h->fp64 && !A.lmix && !A.gtens && red.size() * sizeof(double) <= (64ull << 20)) h->red64 = red;
        const size_t spec_elems = 3ull * A.nzs * A.nkx * A.ly;
        A.spec = dalloc(h, spec_elems * 2 * sizeof(T));
        A.twx = upload_twiddles<T>(h, A.px);
        A.twy = upload_twiddles<T>(h, A.py);
        A.twz = A.pz > 1 ? upload_twiddles<T>(h, A.pz) : A.twy;
        if (!four_step_twiddles<T>(h, A.py, &A.tw4y)) return false;
        // FP32 film at px or py = 1024: three-step FFTs (the y pass's
        // FULLMAG_FDM_YPASS8=0 keeps the 32 x 32 four-step one)
        if (!h->fp64 && nz == 1 && (A.py == 1024 || A.px == 1024) && !f8_bases<T>(h, &A.tw8)) return false;
        {
            const char *e8 = std::getenv("FULLMAG_FDM_YPASS8");
            A.yf8_allowed = !(e8 && e8[0] == '0');
        }
        if (nz > 1 && !four_step_twiddles<T>(h, A.py, &A.tw4y_f, false, true)) return false;
        if (!four_step_twiddles<T>(h, A.px, &A.tw4x)) return false;
        if (nz > 1 && !four_step_twiddles<T>(h, A.pz, &A.tw4z, true)) return false;
        if (!A.spec || !A.twx || !A.twy || !A.twz) return hip_ok(h, hipErrorOutOfMemory, "alloc spectrum");
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) != hipSuccess || lds <= 0)
            lds = 65536;
        h->lds_limit = (size_t)lds - 1024;
        int rows_req = 0, xfast_req = -1, yl_req = 0, zf_req = 0;
        tuned_lookup(h->fp64, nx, ny, nz, A.method, tune_feat(A), &rows_req, &xfast_req, &yl_req, &zf_req);
        A.yl_threads = yl_req;
        A.zf_ub = zf_req;
        // FULLMAG_FDM_YZ_TILE="threads,ky" forces the 3-D y-line and z-pass
        // tiles (invalid values fall back to the defaults in the launches)
        if (const char *yz = std::getenv("FULLMAG_FDM_YZ_TILE")) {
            int y = 0, z = 0;
            if (std::sscanf(yz, "%d,%d", &y, &z) == 2) {
                A.yl_threads = y;
                A.zf_ub = z;
            }
        }
        const char *ce = fm::configure(A, h->lds_limit, rows_req, xfast_req);
        if (rows_req && ce && *ce) ce = fm::configure(A, h->lds_limit);
        if (ce && *ce) {
            h->err = ce;
            return false;
        }
    }
    // partials for the smallest x tile (one row per work-group), so that the
    // tuner may change the tile later
    const long long nparts = std::max<long long>(A.demag ? A.nrows : 0, fm::num_partials(A));
    A.partials = static_cast<double *>(dalloc(h, (size_t)nparts * fm::kStatSlots * sizeof(double)));
    if (!A.partials) return hip_ok(h, hipErrorOutOfMemory, "alloc partials");
    for (int i = 0; i < 3; ++i)
        if (!(h->args_d[i] = dalloc(h, sizeof(fm::Args<T>)))) return hip_ok(h, hipErrorOutOfMemory, "alloc arguments");
    if (!put_args(h, A, 0)) return false;
    return upload_state<T>(h, p->initial_magnetization_xyz, nullptr, 0);
}

bool create_impl(Backend *h, const fullmag_fdm_plan_desc *p, const fullmag_fdm_plan_ext *x) {
    if (select_device(h) != FULLMAG_FDM_OK) return false;
    const auto &g = p->grid;
    if (g.nx == 0 || g.ny == 0 || g.nz == 0) return (h->err = "grid dimensions must be positive", false);
    if (!(g.dx > 0) || !(g.dy > 0) || !(g.dz > 0)) return (h->err = "cell sizes must be positive", false);
    if (!(p->material.saturation_magnetisation > 0)) return (h->err = "saturation magnetisation must be positive", false);
    if (p->precision != FULLMAG_FDM_PRECISION_SINGLE && p->precision != FULLMAG_FDM_PRECISION_DOUBLE)
        return (h->err = "invalid precision", false);
    if (p->integrator != FULLMAG_FDM_INTEGRATOR_DP45 && p->integrator != FULLMAG_FDM_INTEGRATOR_RK4 &&
        p->integrator != FULLMAG_FDM_INTEGRATOR_HEUN && p->integrator != FULLMAG_FDM_INTEGRATOR_RK23 &&
        p->integrator != FULLMAG_FDM_INTEGRATOR_ABM3)
        return (h->err = "integrator not supported by fdm_gpu_native (supported: heun, dp45, rk4, rk23, abm3)", false);
    h->nx = g.nx;
    h->ny = g.ny;
    h->nz = g.nz;
    h->n = (uint64_t)g.nx * g.ny * g.nz;
    // per-vector offsets (3 n) and half-spectrum offsets (~6 n complex) are 32-bit in the kernels
    if ((uint64_t)(g.nx + 1) * g.ny * g.nz * 8 >= (1ull << 31))
        return (h->err = "grid too large: (nx+1)*ny*nz must stay below 2^28 cells", false);
    if (!p->initial_magnetization_xyz || p->initial_magnetization_len != 3 * h->n)
        return (h->err = "initial_magnetization_len mismatch: expected " + std::to_string(3 * h->n) + ", got " +
                         std::to_string(p->initial_magnetization_len),
                false);
    h->fp64 = p->precision == FULLMAG_FDM_PRECISION_DOUBLE;
    if (!hip_ok(h, h->fp64 ? fm::upload_methods<double>() : fm::upload_methods<float>(), "method table"))
        return false;
    if (!hip_ok(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "stream") ||
        !hip_ok(h, hipStreamCreateWithFlags(&h->io_stream, hipStreamNonBlocking), "readout stream"))
        return false;
    if (!hip_ok(h, hipHostMalloc((void **)&h->ctl_h, sizeof(fm::Ctl), hipHostMallocDefault), "pinned ctl"))
        return false;
    if (x && x->struct_size < offsetof(fullmag_fdm_plan_ext, thermal_dt))
        return (h->err = "fullmag_fdm_plan_ext.struct_size too small", false);
    const bool ok = h->fp64 ? create_typed<double>(h, p, x) : create_typed<float>(h, p, x);
    if (!ok) return false;
    std::memset(h->ctl_h, 0, sizeof(fm::Ctl));
    h->ctl_h->dt = 1e-13;
    h->ctl_h->k1_redo = 1;
    if (!push_ctl(h)) return false;
    if (!eval_current(h, true)) return false;
    h->ok = true;
    return true;
}

template <class T>
bool download_soa(Backend *h, const void *dev, std::vector<T> &out) {
    out.resize(3 * h->n);
    return hip_ok(h, hipMemcpy(out.data(), dev, out.size() * sizeof(T), hipMemcpyDeviceToHost), "download");
}

// Waveform factor f(t) of the time-dependent Zeeman term, as the kernels
// evaluate it (FULLMAG_FDM_WAVE_*).
double wave_factor_host(int kind, double t0, double t1, double freq, double phase, double t) {
    constexpr double kPi = 3.14159265358979323846;
    switch (kind) {
        case fm::WAVE_RAMP: {
            const double u = (t - t0) / (t1 - t0);
            return u < 0.0 ? 0.0 : u > 1.0 ? 1.0 : u;
        }
        case fm::WAVE_SINC: {
            const double x = 2.0 * kPi * freq * (t - t0);
            return x == 0.0 ? 1.0 : std::sin(x) / x;
        }
        case fm::WAVE_SINE: return std::sin(2.0 * kPi * freq * t + phase);
        default: return 0.0;
    }
}

// Field as AoS double on the host.
bool field_aos(Backend *h, fullmag_fdm_observable obs, std::vector<double> &out) {
    const uint64_t n = h->n;
    out.assign(3 * n, 0.0);
    if (obs == FULLMAG_FDM_OBSERVABLE_H_EXT) {
        // H_ext(t) = H0 + f(t) H1 (H1 on the waveform's cells) at the time of
        // the current state, the field the kernels add and price in the
        // Zeeman energy
        double h1[3] = {0, 0, 0};
        const double f = with_args(h, [&](auto &A) {
            for (int c = 0; c < 3; ++c) h1[c] = A.wave_amp[c];
            return wave_factor_host(A.wave_kind, A.wave_t0, A.wave_t1, A.wave_freq, A.wave_phase, h->ctl_h->t);
        });
        for (uint64_t i = 0; i < n; ++i) {
            if (h->has_mask && !h->mask_h[i]) continue;
            const bool wave = f != 0.0 && (h->wave_mask_h.empty() || h->wave_mask_h[i]);
            for (int c = 0; c < 3; ++c) out[3 * i + c] = (h->has_ext ? h->hext[c] : 0.0) + (wave ? f * h1[c] : 0.0);
        }
        return true;
    }
    const void *src = nullptr;
    if (obs == FULLMAG_FDM_OBSERVABLE_M) {
        src = with_args(h, [&](auto &A) -> const void * { return A.m_all + (size_t)h->ctl_h->parity * 3 * n; });
    } else {
        if (!h->obs_fresh && !eval_current(h, true)) return false;
        const int w = obs == FULLMAG_FDM_OBSERVABLE_H_EX ? 0 : obs == FULLMAG_FDM_OBSERVABLE_H_DEMAG ? 1 : 2;
        src = h->obs_base[w];
    }
    if (h->fp64) {
        std::vector<double> s;
        if (!download_soa(h, src, s)) return false;
        for (uint64_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) out[3 * i + c] = s[c * n + i];
    } else {
        std::vector<float> s;
        if (!download_soa(h, src, s)) return false;
        for (uint64_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) out[3 * i + c] = (double)s[c * n + i];
    }
    return true;
}

bool valid_observable(int o) { return o >= FULLMAG_FDM_OBSERVABLE_M && o <= FULLMAG_FDM_OBSERVABLE_H_EFF; }

// Preview bins of H_ext (reference context.cu:1530-1600: the external field
// is not a device field, so its preview is formed on the host): the mean over
// each bin (the bin rule of context.cu:254-317) of H0 on magnetic cells plus
// f(t) H1 on the waveform's cells, FP64 sums in z, y, x order, AoS output.
// Writes only `out` (no allocation).
void hext_bins(Backend *h, uint32_t pnx, uint32_t pny, uint32_t pnz, uint32_t z0, uint32_t zs, double *out) {
    const uint32_t fx = h->nx, fy = h->ny, fz = h->nz;
    double h1[3] = {0, 0, 0};
    const double f = with_args(h, [&](auto &A) {
        for (int c = 0; c < 3; ++c) h1[c] = A.wave_amp[c];
        return wave_factor_host(A.wave_kind, A.wave_t0, A.wave_t1, A.wave_freq, A.wave_phase, h->ctl_h->t);
    });
    auto range = [](uint32_t p, uint32_t full, uint32_t np, uint32_t &b, uint32_t &e) {
        b = (uint32_t)((uint64_t)p * full / np);
        e = (uint32_t)((uint64_t)(p + 1) * full / np);
        if (e <= b) e = b + 1;
        if (e > full) e = full;
    };
    for (uint32_t pz = 0; pz < pnz; ++pz)
        for (uint32_t py = 0; py < pny; ++py)
            for (uint32_t px = 0; px < pnx; ++px) {
                uint32_t xs, xe, ys, ye;
                range(px, fx, pnx, xs, xe);
                range(py, fy, pny, ys, ye);
                uint32_t zs0 = z0 + pz * zs;
                if (zs0 >= fz) zs0 = fz - 1;
                uint32_t ze = z0 + (pz + 1) * zs;
                if (ze <= zs0) ze = zs0 + 1;
                if (ze > fz) ze = fz;
                double a[3] = {0, 0, 0}, cnt = 0;
                for (uint32_t z = zs0; z < ze; ++z)
                    for (uint32_t y = ys; y < ye; ++y)
                        for (uint32_t x = xs; x < xe; ++x) {
                            const size_t i = ((size_t)z * fy + y) * fx + x;
                            cnt += 1.0;
                            if (h->has_mask && !h->mask_h[i]) continue;
                            const bool wave = f != 0.0 && (h->wave_mask_h.empty() || h->wave_mask_h[i]);
                            for (int c = 0; c < 3; ++c)
                                a[c] += (h->has_ext ? h->hext[c] : 0.0) + (wave ? f * h1[c] : 0.0);
                        }
                const size_t o = ((size_t)pz * pny + py) * pnx + px;
                for (int c = 0; c < 3; ++c) out[3 * o + c] = a[c] / cnt;
            }
}

}  // namespace

extern "C" {

int fullmag_fdm_is_available(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx9", 4) == 0 ? 1 : 0;
}

fullmag_fdm_backend *fullmag_fdm_backend_create(const fullmag_fdm_plan_desc *plan) {
    return fullmag_fdm_backend_create_ext(plan, nullptr);
}

fullmag_fdm_backend *fullmag_fdm_backend_create_ext(const fullmag_fdm_plan_desc *plan,
                                                    const fullmag_fdm_plan_ext *ext) {
    if (!plan) return nullptr;
    Backend *h = new (std::nothrow) Backend();
    if (!h) return nullptr;
    try {
        if (!create_impl(h, plan, ext) && h->err.empty()) h->err = "create failed";
    } catch (const std::bad_alloc &) {
        h->err = "host allocation failed";
    } catch (...) {
        h->err = "internal error during create";
    }
    return h;
}

int fullmag_fdm_backend_step(fullmag_fdm_backend *h, double dt, fullmag_fdm_step_stats *out) {
    if (!h) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, h->err.empty() ? "backend not initialised" : h->err);
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(h, FULLMAG_FDM_ERR_INVALID, "dt must be positive and finite");
    h->err.clear();
    hipSetDevice(h->device);
    const auto t0 = std::chrono::steady_clock::now();
    const double dmin = h->fp64 ? h->ad.dt_min : h->af.dt_min, dmax = h->fp64 ? h->ad.dt_max : h->af.dt_max;
    h->ctl_h->dt = h->adaptive ? std::min(std::max(dt, dmin), dmax) : dt;
    h->ctl_h->loop = 0;
    h->ctl_h->done = 0;
    if (!push_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    if (!h->poll) ensure_graph(h);
    h->obs_fresh = false;
    for (;;) {
        if (h->poll && h->poll(h->poll_user)) {
            fill_stats(h, out, 0.0, 0);
            return fail(h, FULLMAG_FDM_ERR_INTERRUPTED, "interrupted");
        }
        if (h->poll) {
            // stage-by-stage with a poll between stages; the state is only
            // replaced by the controller, so an interrupt leaves m untouched
            bool interrupted = false;
            const bool ok = with_args(h, [&](auto &A) {
                if (!hip_ok(h, fm::enqueue_k1(A, h->stream), "launch (k1)")) return false;
                for (int s = 2; s < h->nstage; ++s) {
                    if (!hip_ok(h, fm::enqueue_eval(A, s, 0, h->stream), "launch")) return false;
                    if (h->poll(h->poll_user)) {
                        interrupted = true;
                        return true;
                    }
                }
                return hip_ok(h, fm::enqueue_eval(A, h->nstage, 1, h->stream), "launch (final stage)");
            });
            if (!ok) return FULLMAG_FDM_ERR_CUDA;
            if (interrupted) {
                hipStreamSynchronize(h->stream);
                fill_stats(h, out, 0.0, 0);
                return fail(h, FULLMAG_FDM_ERR_INTERRUPTED, "interrupted");
            }
        } else if (!launch_attempt(h)) {
            return FULLMAG_FDM_ERR_CUDA;
        }
        if (!pull_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
        if (h->ctl_h->accepted) break;
    }
    const uint64_t ns =
        (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    fill_stats(h, out, h->ctl_h->dt_used, ns);
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_set_interrupt_poll(fullmag_fdm_backend *h, fullmag_fdm_interrupt_poll_fn fn, void *user) {
    if (!h) return FULLMAG_FDM_ERR_INVALID;
    h->poll = fn;
    h->poll_user = user;
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_copy_field_f64(fullmag_fdm_backend *h, fullmag_fdm_observable obs, double *out,
                                       uint64_t len) {
    if (!h || !out) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (!valid_observable(obs)) return fail(h, FULLMAG_FDM_ERR_INVALID, "unknown observable");
    if (len != 3 * h->n) return fail(h, FULLMAG_FDM_ERR_INVALID, "out_len mismatch");
    hipSetDevice(h->device);
    std::vector<double> f;
    if (!field_aos(h, obs, f)) return FULLMAG_FDM_ERR_CUDA;
    std::memcpy(out, f.data(), f.size() * sizeof(double));
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_copy_field_f32(fullmag_fdm_backend *h, fullmag_fdm_observable obs, float *out, uint64_t len) {
    if (!h || !out) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (!valid_observable(obs)) return fail(h, FULLMAG_FDM_ERR_INVALID, "unknown observable");
    if (len != 3 * h->n) return fail(h, FULLMAG_FDM_ERR_INVALID, "out_len mismatch");
    hipSetDevice(h->device);
    std::vector<double> f;
    if (!field_aos(h, obs, f)) return FULLMAG_FDM_ERR_CUDA;
    for (size_t i = 0; i < f.size(); ++i) out[i] = (float)f[i];
    return FULLMAG_FDM_OK;
}

// Device address of an observable's SoA field in the state precision, the
// fields evaluated at the current state first if needed.  H_EXT has no
// device copy (it is built on the host).
static bool field_device_src(fullmag_fdm_backend *h, fullmag_fdm_observable obs, const void **src) {
    if (obs == FULLMAG_FDM_OBSERVABLE_M) {
        *src = with_args(h, [&](auto &A) -> const void * { return A.m_all + (size_t)h->ctl_h->parity * 3 * h->n; });
        return true;
    }
    if (!h->obs_fresh && !eval_current(h, true)) return false;
    *src = h->obs_base[obs == FULLMAG_FDM_OBSERVABLE_H_EX ? 0 : obs == FULLMAG_FDM_OBSERVABLE_H_DEMAG ? 1 : 2];
    return true;
}

// Preview bins of a device field into device memory (AoS, FP64 or FP32), on
// the compute stream.
static bool preview_launch(fullmag_fdm_backend *h, const void *src, const fm::PreviewGeom &g, void *out, int out_fp64,
                           hipStream_t st) {
    const hipError_t e =
        h->fp64 ? fm::enqueue_preview<double>(static_cast<const double *>(src), (long long)h->n, g, out, out_fp64, st)
                : fm::enqueue_preview<float>(static_cast<const float *>(src), (long long)h->n, g, out, out_fp64, st);
    return hip_ok(h, e, "preview binning");
}

static int preview_check(fullmag_fdm_backend *h, fullmag_fdm_observable obs, uint32_t pnx, uint32_t pny, uint32_t pnz,
                         uint32_t zs) {
    if (!h) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (!valid_observable(obs) || pnx == 0 || pny == 0 || pnz == 0 || zs == 0)
        return fail(h, FULLMAG_FDM_ERR_INVALID, "invalid preview request");
    return FULLMAG_FDM_OK;
}

// Synchronous preview (copy_field_preview_*): binned on the device, only the
// bins cross to the host.  H_EXT is binned on the host from its field.
static int preview_copy(fullmag_fdm_backend *h, fullmag_fdm_observable obs, uint32_t pnx, uint32_t pny, uint32_t pnz,
                        uint32_t z0, uint32_t zs, void *out, uint64_t len, bool out_fp64) {
    const int rc = preview_check(h, obs, pnx, pny, pnz, zs);
    if (rc != FULLMAG_FDM_OK) return rc;
    if (!out) return fail(h, FULLMAG_FDM_ERR_INVALID, "preview output is NULL");
    const uint64_t cnt = (uint64_t)pnx * pny * pnz;
    if (len != cnt * 3) return fail(h, FULLMAG_FDM_ERR_INVALID, "preview out_len mismatch");
    hipSetDevice(h->device);
    if (obs == FULLMAG_FDM_OBSERVABLE_H_EXT) {
        if (out_fp64) {
            hext_bins(h, pnx, pny, pnz, z0, zs, static_cast<double *>(out));
        } else {
            std::vector<double> v(len);
            hext_bins(h, pnx, pny, pnz, z0, zs, v.data());
            for (size_t i = 0; i < v.size(); ++i) static_cast<float *>(out)[i] = (float)v[i];
        }
        return FULLMAG_FDM_OK;
    }
    const void *src = nullptr;
    if (!field_device_src(h, obs, &src)) return FULLMAG_FDM_ERR_CUDA;
    const size_t bytes = (size_t)len * (out_fp64 ? sizeof(double) : sizeof(float));
    if (bytes > h->preview_cap) {
        if (!hip_ok(h, hipStreamSynchronize(h->stream), "stream sync")) return FULLMAG_FDM_ERR_CUDA;
        if (h->preview_d) hipFree(h->preview_d);
        h->preview_d = nullptr;
        h->preview_cap = 0;
        if (!hip_ok(h, hipMalloc(&h->preview_d, bytes), "preview buffer")) return FULLMAG_FDM_ERR_CUDA;
        h->preview_cap = bytes;
    }
    const fm::PreviewGeom g{h->nx, h->ny, h->nz, pnx, pny, pnz, z0, zs};
    if (!preview_launch(h, src, g, h->preview_d, out_fp64 ? 1 : 0, h->stream) ||
        !hip_ok(h, hipMemcpyAsync(out, h->preview_d, bytes, hipMemcpyDeviceToHost, h->stream), "preview download") ||
        !hip_ok(h, hipStreamSynchronize(h->stream), "stream sync"))
        return FULLMAG_FDM_ERR_CUDA;
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_copy_field_preview_f64(fullmag_fdm_backend *h, fullmag_fdm_observable obs, uint32_t pnx,
                                               uint32_t pny, uint32_t pnz, uint32_t z0, uint32_t zs, double *out,
                                               uint64_t len) {
    return preview_copy(h, obs, pnx, pny, pnz, z0, zs, out, len, true);
}

int fullmag_fdm_backend_copy_field_preview_f32(fullmag_fdm_backend *h, fullmag_fdm_observable obs, uint32_t pnx,
                                               uint32_t pny, uint32_t pnz, uint32_t z0, uint32_t zs, float *out,
                                               uint64_t len) {
    return preview_copy(h, obs, pnx, pny, pnz, z0, zs, out, len, false);
}

// copy_field_preview_f64 without any device allocation: the bins are staged
// in the library's fixed per-GPU array (kPreviewScratchBytes) instead of the
// handle's grown-on-demand buffer.  Same kernel, same result.
int fullmag_fdm_backend_copy_field_preview_f64_noalloc(fullmag_fdm_backend *h, fullmag_fdm_observable obs,
                                                       uint32_t pnx, uint32_t pny, uint32_t pnz, uint32_t z0,
                                                       uint32_t zs, double *out, uint64_t len) {
    const int rc = preview_check(h, obs, pnx, pny, pnz, zs);
    if (rc != FULLMAG_FDM_OK) return rc;
    if (!out) return fail(h, FULLMAG_FDM_ERR_INVALID, "preview output is NULL");
    const uint64_t cnt = (uint64_t)pnx * pny * pnz;
    if (len != cnt * 3) return fail(h, FULLMAG_FDM_ERR_INVALID, "preview out_len mismatch");
    const uint64_t bytes = len * sizeof(double);
    if (bytes > fm::kPreviewScratchBytes)
        return fail(h, FULLMAG_FDM_ERR_INVALID,
                    "preview of " + std::to_string(cnt) + " cells exceeds the allocation-free staging array (" +
                        std::to_string(fm::kPreviewScratchBytes / (3 * sizeof(double))) + " cells)");
    hipSetDevice(h->device);
    if (obs == FULLMAG_FDM_OBSERVABLE_H_EXT) {
        hext_bins(h, pnx, pny, pnz, z0, zs, out);
        return FULLMAG_FDM_OK;
    }
    void *scr = nullptr;
    if (!hip_ok(h, fm::preview_scratch(h->device, &scr), "preview staging array")) return FULLMAG_FDM_ERR_CUDA;
    const void *src = nullptr;
    if (!field_device_src(h, obs, &src)) return FULLMAG_FDM_ERR_CUDA;
    const fm::PreviewGeom g{h->nx, h->ny, h->nz, pnx, pny, pnz, z0, zs};
    if (!preview_launch(h, src, g, scr, 1, h->stream) ||
        !hip_ok(h, hipMemcpyAsync(out, scr, bytes, hipMemcpyDeviceToHost, h->stream), "preview download") ||
        !hip_ok(h, hipStreamSynchronize(h->stream), "stream sync"))
        return FULLMAG_FDM_ERR_CUDA;
    return FULLMAG_FDM_OK;
}

}  // extern "C"

namespace {

// Releases what a snapshot holds (either kind).
template <class S>
void snapshot_release(S *s) {
    if (!s) return;
    if (s->done) hipEventSynchronize(s->done);
    if (s->ready) hipEventDestroy(s->ready);
    if (s->done) hipEventDestroy(s->done);
    if (s->staging) hipFree(s->staging);
    if (s->host) hipHostFree(s->host);
    delete s;
}

// Schedules the staged payload's download: `ready` on the compute stream,
// the copy on the readout stream after it, `done` behind the copy.
template <class S>
bool snapshot_schedule(fullmag_fdm_backend *h, S *s) {
    return hip_ok(h, hipEventCreateWithFlags(&s->ready, hipEventDisableTiming), "snapshot event") &&
           hip_ok(h, hipEventCreateWithFlags(&s->done, hipEventDisableTiming), "snapshot event") &&
           hip_ok(h, hipEventRecord(s->ready, h->stream), "snapshot event") &&
           hip_ok(h, hipStreamWaitEvent(h->io_stream, s->ready, 0), "snapshot stream wait") &&
           hip_ok(h, hipMemcpyAsync(s->host, s->staging, s->bytes, hipMemcpyDeviceToHost, h->io_stream),
                  "snapshot download") &&
           hip_ok(h, hipEventRecord(s->done, h->io_stream), "snapshot event");
}

template <class S>
int snapshot_wait(S *s, const void **data, uint64_t *len, fullmag_fdm_snapshot_desc *desc) {
    if (!s) return FULLMAG_FDM_ERR_INVALID;
    if (s->done && !s->waited) {
        if (hipEventSynchronize(s->done) != hipSuccess) return FULLMAG_FDM_ERR_CUDA;
        s->waited = true;
    }
    if (data) *data = s->host;
    if (len) *len = s->bytes;
    if (desc) *desc = s->desc;
    return FULLMAG_FDM_OK;
}

}  // namespace

extern "C" {

fullmag_fdm_field_snapshot *fullmag_fdm_backend_begin_field_snapshot(fullmag_fdm_backend *h,
                                                                     fullmag_fdm_observable obs) {
    if (!h || !h->ok || !valid_observable(obs)) {
        if (h) h->err = "invalid snapshot request";
        return nullptr;
    }
    hipSetDevice(h->device);
    auto *s = new (std::nothrow) fullmag_fdm_field_snapshot();
    if (!s) return nullptr;
    const size_t sb = h->fp64 ? 8 : 4;
    s->bytes = 3 * h->n * sb;
    s->desc.cell_count = h->n;
    s->desc.component_count = 3;
    s->desc.scalar_bytes = (uint32_t)sb;
    s->desc.scalar_type = h->fp64 ? FULLMAG_FDM_SNAPSHOT_SCALAR_F64 : FULLMAG_FDM_SNAPSHOT_SCALAR_F32;
    if (!hip_ok(h, hipHostMalloc(&s->host, s->bytes, hipHostMallocDefault), "pinned snapshot allocation")) {
        snapshot_release(s);
        return nullptr;
    }
    if (obs == FULLMAG_FDM_OBSERVABLE_H_EXT) {
        // built on the host: complete at once
        std::vector<double> f;
        if (!field_aos(h, obs, f)) {
            snapshot_release(s);
            return nullptr;
        }
        for (uint64_t i = 0; i < h->n; ++i)
            for (int c = 0; c < 3; ++c) {
                if (h->fp64) static_cast<double *>(s->host)[c * h->n + i] = f[3 * i + c];
                else static_cast<float *>(s->host)[c * h->n + i] = (float)f[3 * i + c];
            }
        s->waited = true;
        return s;
    }
    const void *src = nullptr;
    if (!field_device_src(h, obs, &src) || !hip_ok(h, hipMalloc(&s->staging, s->bytes), "snapshot staging") ||
        !hip_ok(h, hipMemcpyAsync(s->staging, src, s->bytes, hipMemcpyDeviceToDevice, h->stream), "snapshot staging") ||
        !snapshot_schedule(h, s)) {
        snapshot_release(s);
        return nullptr;
    }
    return s;
}

fullmag_fdm_preview_snapshot *fullmag_fdm_backend_begin_preview_snapshot(fullmag_fdm_backend *h,
                                                                         fullmag_fdm_observable obs, uint32_t pnx,
                                                                         uint32_t pny, uint32_t pnz, uint32_t z0,
                                                                         uint32_t zs) {
    if (preview_check(h, obs, pnx, pny, pnz, zs) != FULLMAG_FDM_OK) return nullptr;
    hipSetDevice(h->device);
    auto *s = new (std::nothrow) fullmag_fdm_preview_snapshot();
    if (!s) return nullptr;
    const uint64_t cells = (uint64_t)pnx * pny * pnz;
    const size_t sb = h->fp64 ? 8 : 4;
    s->bytes = cells * 3 * sb;
    s->desc.cell_count = cells;
    s->desc.component_count = 3;
    s->desc.scalar_bytes = (uint32_t)sb;
    s->desc.scalar_type = h->fp64 ? FULLMAG_FDM_SNAPSHOT_SCALAR_F64 : FULLMAG_FDM_SNAPSHOT_SCALAR_F32;
    if (!hip_ok(h, hipHostMalloc(&s->host, s->bytes, hipHostMallocDefault), "pinned preview allocation")) {
        snapshot_release(s);
        return nullptr;
    }
    if (obs == FULLMAG_FDM_OBSERVABLE_H_EXT) {
        std::vector<double> v(cells * 3);
        hext_bins(h, pnx, pny, pnz, z0, zs, v.data());
        for (size_t i = 0; i < v.size(); ++i) {
            if (h->fp64) static_cast<double *>(s->host)[i] = v[i];
            else static_cast<float *>(s->host)[i] = (float)v[i];
        }
        s->waited = true;
        return s;
    }
    const void *src = nullptr;
    const fm::PreviewGeom g{h->nx, h->ny, h->nz, pnx, pny, pnz, z0, zs};
    if (!field_device_src(h, obs, &src) || !hip_ok(h, hipMalloc(&s->staging, s->bytes), "preview staging") ||
        !preview_launch(h, src, g, s->staging, h->fp64 ? 1 : 0, h->stream) || !snapshot_schedule(h, s)) {
        snapshot_release(s);
        return nullptr;
    }
    return s;
}

int fullmag_fdm_field_snapshot_wait(fullmag_fdm_field_snapshot *s, const void **data, uint64_t *len,
                                    fullmag_fdm_snapshot_desc *desc) {
    return snapshot_wait(s, data, len, desc);
}

int fullmag_fdm_preview_snapshot_wait(fullmag_fdm_preview_snapshot *s, const void **data, uint64_t *len,
                                      fullmag_fdm_snapshot_desc *desc) {
    return snapshot_wait(s, data, len, desc);
}

void fullmag_fdm_field_snapshot_destroy(fullmag_fdm_field_snapshot *s) { snapshot_release(s); }

void fullmag_fdm_preview_snapshot_destroy(fullmag_fdm_preview_snapshot *s) { snapshot_release(s); }

static int upload_common(fullmag_fdm_backend *h, const double *d, const float *f, uint64_t len) {
    if (!h || (!d && !f)) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (len != 3 * h->n) return fail(h, FULLMAG_FDM_ERR_INVALID, "magnetisation length mismatch");
    hipSetDevice(h->device);
    if (!pull_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    const int slot = h->ctl_h->parity;
    const bool ok = h->fp64 ? upload_state<double>(h, d, f, slot) : upload_state<float>(h, d, f, slot);
    if (!ok) return FULLMAG_FDM_ERR_CUDA;
    if (h->multistep) {
        // external state change: the RHS history no longer belongs to the
        // trajectory (ExchangeLlgState::reset_abm_history, fdm_state.rs:172)
        h->ctl_h->abm_count = 0;
        h->ctl_h->abm_dt = 0.0;
    }
    // nor does the FSAL stage (k_fsal is dropped with the state)
    h->ctl_h->k1_redo = 1;
    if (!push_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    // the first stage of the next step needs k1 of the new state
    if (!eval_current(h, true)) return FULLMAG_FDM_ERR_CUDA;
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_upload_magnetization_f64(fullmag_fdm_backend *h, const double *m, uint64_t len) {
    return upload_common(h, m, nullptr, len);
}

int fullmag_fdm_backend_upload_magnetization_f32(fullmag_fdm_backend *h, const float *m, uint64_t len) {
    return upload_common(h, nullptr, m, len);
}

int fullmag_fdm_backend_refresh_observables(fullmag_fdm_backend *h) {
    if (!h) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    hipSetDevice(h->device);
    return eval_current(h, true) ? FULLMAG_FDM_OK : FULLMAG_FDM_ERR_CUDA;
}

int fullmag_fdm_backend_refresh_demag_observable(fullmag_fdm_backend *h) {
    return fullmag_fdm_backend_refresh_observables(h);
}

int fullmag_fdm_backend_snapshot_stats(fullmag_fdm_backend *h, fullmag_fdm_step_stats *out) {
    if (!h || !out) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    hipSetDevice(h->device);
    if (!eval_current(h, true)) return FULLMAG_FDM_ERR_CUDA;
    fill_stats(h, out, h->ctl_h->dt_used, 0);
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_get_device_info(fullmag_fdm_backend *h, fullmag_fdm_device_info *info) {
    if (!h || !info) return FULLMAG_FDM_ERR_INVALID;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) return fail(h, FULLMAG_FDM_ERR_CUDA, "device query");
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->name, sizeof(info->name), "%s (%s)", prop.name, prop.gcnArchName);
    info->compute_capability_major = prop.major;
    info->compute_capability_minor = prop.minor;
    int v = 0;
    if (hipDriverGetVersion(&v) == hipSuccess) info->driver_version = v;
    if (hipRuntimeGetVersion(&v) == hipSuccess) info->runtime_version = v;
    return FULLMAG_FDM_OK;
}

const char *fullmag_fdm_backend_last_error(fullmag_fdm_backend *h) {
    if (!h || h->err.empty()) return nullptr;
    return h->err.c_str();
}

void fullmag_fdm_backend_destroy(fullmag_fdm_backend *h) {
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
    if (h->graph_chunk) hipGraphExecDestroy(h->graph_chunk);
    if (h->graph) hipGraphDestroy(h->graph);
    for (void *p : h->allocs) hipFree(p);
    for (void *p : h->adj.bufs) hipFree(p);
    if (h->io_stream) hipStreamSynchronize(h->io_stream);
    if (h->preview_d) hipFree(h->preview_d);
    if (h->ctl_h) hipHostFree(h->ctl_h);
    if (h->stream) hipStreamDestroy(h->stream);
    if (h->io_stream) hipStreamDestroy(h->io_stream);
    delete h;
}

// ------------------------------------------------------------ extensions --

int fullmag_fdm_ext_version(void) { return FULLMAG_FDM_EXT_VERSION; }

int fullmag_fdm_random_unit_vectors(uint64_t seed, uint64_t count, double *out) {
    if (seed == 0 || (!out && count)) return FULLMAG_FDM_ERR_INVALID;
    uint64_t st = seed;
    const double scale = (double)UINT64_MAX;
    for (uint64_t i = 0; i < count; ++i) {
        for (;;) {
            double c[3];
            for (int k = 0; k < 3; ++k) {
                st ^= st << 13;
                st ^= st >> 7;
                st ^= st << 17;
                c[k] = ((double)st / scale) * 2.0 - 1.0;
            }
            const double nrm = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            if (nrm > 1e-10) {
                for (int k = 0; k < 3; ++k) out[3 * i + k] = c[k] / nrm;
                break;
            }
        }
    }
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_advance(fullmag_fdm_backend *h, double t_end, uint64_t max_steps, double dt,
                                fullmag_fdm_step_stats *out, uint64_t *out_steps) {
    if (!h) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (!(dt > 0.0) || !std::isfinite(dt) || std::isnan(t_end))
        return fail(h, FULLMAG_FDM_ERR_INVALID, "dt must be positive and finite");
    h->err.clear();
    hipSetDevice(h->device);
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t step0 = h->ctl_h->step;
    fm::Ctl &c = *h->ctl_h;
    const double dmin = h->fp64 ? h->ad.dt_min : h->af.dt_min, dmax = h->fp64 ? h->ad.dt_max : h->af.dt_max;
    c.loop = 1;
    c.t_end = t_end;
    c.step_end = sat_add(c.step, max_steps);
    c.dt_carry = dt;
    double d = std::min(dt, t_end - c.t);
    if (h->adaptive) d = std::min(std::max(d, dmin), dmax);
    c.dt = d;
    c.done = (c.t >= t_end || max_steps == 0) ? 1 : 0;
    if (!push_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    ensure_graph(h);
    h->obs_fresh = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    fm::Ctl *flag_h = nullptr;
    if (hipHostMalloc((void **)&flag_h, 2 * sizeof(fm::Ctl), hipHostMallocDefault) != hipSuccess)
        return fail(h, FULLMAG_FDM_ERR_CUDA, "pinned alloc");
    hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    int rc = FULLMAG_FDM_OK;
    // Trial steps are queued in chunks; chunk i+1 is queued before chunk i's
    // control block is read back.  Without rejections exactly max_steps
    // trials are queued (no early-exit launches); once every planned trial is
    // queued the host waits and tops up by the number of steps still missing.
    uint64_t launched = 0, planned = max_steps;
    int cur = 0;
    bool pending = false;
    while (!c.done) {
        uint64_t nl = kAttemptChunk;
        if (launched < planned) {
            nl = std::min<uint64_t>(nl, planned - launched);
        } else {
            if (hipStreamSynchronize(h->stream) != hipSuccess ||
                hipMemcpy(&flag_h[cur], h->ctl_d, sizeof(fm::Ctl), hipMemcpyDeviceToHost) != hipSuccess) {
                rc = fail(h, FULLMAG_FDM_ERR_CUDA, "control readback");
                break;
            }
            pending = false;
            if (flag_h[cur].done) break;
            const uint64_t missing = flag_h[cur].step_end > flag_h[cur].step ? flag_h[cur].step_end - flag_h[cur].step : 1;
            nl = std::min<uint64_t>(kAttemptChunk, missing);
            planned = launched + nl;
        }
        if (nl == (uint64_t)kAttemptChunk && h->graph_chunk) {
            if (!hip_ok(h, hipGraphLaunch(h->graph_chunk, h->stream), "graph launch")) rc = FULLMAG_FDM_ERR_CUDA;
        } else {
            for (uint64_t a = 0; a < nl; ++a)
                if (!launch_attempt(h)) {
                    rc = FULLMAG_FDM_ERR_CUDA;
                    break;
                }
        }
        if (rc != FULLMAG_FDM_OK) break;
        launched += nl;
        if (!hip_ok(h, hipMemcpyAsync(&flag_h[cur], h->ctl_d, sizeof(fm::Ctl), hipMemcpyDeviceToHost, h->stream),
                    "control readback") ||
            !hip_ok(h, hipEventRecord(ev[cur], h->stream), "event record")) {
            rc = FULLMAG_FDM_ERR_CUDA;
            break;
        }
        if (pending) {
            if (hipEventSynchronize(ev[1 - cur]) != hipSuccess) {
                rc = fail(h, FULLMAG_FDM_ERR_CUDA, "event sync");
                break;
            }
            if (flag_h[1 - cur].done) break;
        }
        pending = true;
        cur = 1 - cur;
        if (h->poll && h->poll(h->poll_user)) {
            rc = FULLMAG_FDM_ERR_INTERRUPTED;
            break;
        }
    }
    hipStreamSynchronize(h->stream);
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    hipHostFree(flag_h);
    if (rc == FULLMAG_FDM_ERR_INTERRUPTED) {
        // stop the device loop where it is: the state is the last accepted one
        if (!pull_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    } else if (rc != FULLMAG_FDM_OK) {
        return rc;
    } else if (!pull_ctl(h)) {
        return FULLMAG_FDM_ERR_CUDA;
    }
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(h, FULLMAG_FDM_ERR_CUDA, std::string("device error: ") + hipGetErrorString(le));
    c.loop = 0;
    c.done = 0;
    if (!push_ctl(h) || !pull_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    const uint64_t ns =
        (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    fill_stats(h, out, c.dt_used, ns);
    if (out_steps) *out_steps = c.step - step0;
    if (rc == FULLMAG_FDM_ERR_INTERRUPTED) return fail(h, rc, "interrupted");
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_set_adaptive_ext(fullmag_fdm_backend *h, double rtol, double growth, double shrink) {
    if (!h) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (rtol < 0 || shrink < 0 || std::isnan(growth)) return fail(h, FULLMAG_FDM_ERR_INVALID, "invalid tolerances");
    const double g = growth <= 0 ? std::numeric_limits<double>::infinity() : growth;
    h->af.rtol = h->ad.rtol = rtol;
    h->af.growth = h->ad.growth = g;
    h->af.shrink = h->ad.shrink = shrink;
    const bool put = h->fp64 ? put_args(h, h->ad, 0) : put_args(h, h->af, 0);
    if (!put) return FULLMAG_FDM_ERR_CUDA;
    // the graphs read the kernel arguments from device memory (updated above)
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_backend_get_counters(fullmag_fdm_backend *h, uint64_t *acc, uint64_t *rej, double *t) {
    if (!h) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (acc) *acc = h->ctl_h->step;
    if (rej) *rej = h->ctl_h->rejected;
    if (t) *t = h->ctl_h->t;
    return FULLMAG_FDM_OK;
}

namespace {

// The timing launches of a trial stage write that stage's RHS; with ABM3 in
// its predictor-corrector phase that is a slot of the RHS history, so tuning
// calls keep a copy of the ring and put it back when they end.
struct HistoryGuard {
    Backend *h;
    void *buf = nullptr, *ring = nullptr;
    size_t bytes = 0;
    explicit HistoryGuard(Backend *b) : h(b) {
        if (!h->multistep) return;
        const size_t es = h->fp64 ? sizeof(double) : sizeof(float);
        ring = h->fp64 ? (void *)(h->ad.k_all + (size_t)fm::kHistSlot * 3 * h->n)
                       : (void *)(h->af.k_all + (size_t)fm::kHistSlot * 3 * h->n);
        bytes = 3 * 3 * h->n * es;
        if (hipMalloc(&buf, bytes) != hipSuccess) {
            buf = nullptr;
            return;
        }
        hipMemcpyAsync(buf, ring, bytes, hipMemcpyDeviceToDevice, h->stream);
    }
    bool ok() const { return !h->multistep || buf; }
    ~HistoryGuard() {
        if (!buf) return;
        hipMemcpyAsync(ring, buf, bytes, hipMemcpyDeviceToDevice, h->stream);
        hipStreamSynchronize(h->stream);
        hipFree(buf);
    }
};

}  // namespace

int fullmag_fdm_backend_time_kernels(fullmag_fdm_backend *h, int reps, double *out_us, int n_out) {
    if (!h || !out_us || n_out < 7 || reps < 1) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    hipSetDevice(h->device);
    for (int i = 0; i < n_out; ++i) out_us[i] = 0.0;
    if (!pull_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    // The kernels of stage 2 only write trial buffers (stage state, stage-2
    // RHS, spectrum, partials); the controller may change dt or parity, so the
    // control block is restored after every kind.  m and k1 are never written.
    const fm::Ctl saved = *h->ctl_h;
    HistoryGuard keep(h);
    if (!keep.ok()) return fail(h, FULLMAG_FDM_ERR_CUDA, "history copy allocation failed");
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    int rc = FULLMAG_FDM_OK;
    for (int kind = 0; kind < 7 && rc == FULLMAG_FDM_OK; ++kind) {
        if (with_args(h, [&](auto &A) { return fm::enqueue_kind(A, kind, 2, h->stream); }) != hipSuccess) continue;
        hipEventRecord(e0, h->stream);
        for (int r = 0; r < reps; ++r) with_args(h, [&](auto &A) { return fm::enqueue_kind(A, kind, 2, h->stream); });
        hipEventRecord(e1, h->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(h, FULLMAG_FDM_ERR_CUDA, "timing sync");
            break;
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        out_us[kind] = 1000.0 * ms / reps;
        // the controller (kind 5) accepts and flips the state parity; the
        // next kind must see the original slots, or a joined launch into
        // the final stage would write the current state
        *h->ctl_h = saved;
        if (!push_ctl(h)) rc = FULLMAG_FDM_ERR_CUDA;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *h->ctl_h = saved;
    if (!push_ctl(h) || !pull_ctl(h)) rc = FULLMAG_FDM_ERR_CUDA;
    h->obs_fresh = false;
    return rc;
}

int fullmag_fdm_backend_time_trial_kernels(fullmag_fdm_backend *h, int trials, double *out_us, int n_out) {
    if (!h || !out_us || n_out < 7 || trials < 1) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    hipSetDevice(h->device);
    for (int i = 0; i < n_out; ++i) out_us[i] = 0.0;
    const bool demag = with_args(h, [&](auto &A) { return A.demag != 0 && !A.lmix; });
    if (!demag) return fail(h, FULLMAG_FDM_ERR_INVALID, "trial-step timing needs a demag plan without layer stacks");
    if (!pull_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    // the launches of a trial step, as the captured graph has them:
    // (kind, stage)
    const bool film = h->nz == 1;
    const int ns = h->nstage;
    std::vector<std::pair<int, int>> seq{{0, 2}};
    for (int s = 2; s <= ns; ++s) {
        seq.push_back({1, s});
        if (!film) {
            seq.push_back({2, s});
            seq.push_back({3, s});
        }
        seq.push_back({s < ns ? 6 : 4, s});
    }
    seq.push_back({5, ns});
    // a trial step leaves m, k1 (and with the history guard the ABM3 ring)
    // as they were; the controller's decisions live in the control block,
    // which is restored after each trial step
    const fm::Ctl saved = *h->ctl_h;
    HistoryGuard keep(h);
    if (!keep.ok()) return fail(h, FULLMAG_FDM_ERR_CUDA, "history copy allocation failed");
    std::vector<hipEvent_t> ev(seq.size() + 1);
    for (auto &e : ev) hipEventCreate(&e);
    double sum[7] = {0.0};
    int cnt[7] = {0};
    double with_events = 0.0, plain = 0.0;
    int rc = FULLMAG_FDM_OK;
    // trials with an event after every launch, then as many with events
    // only around the whole trial step: the difference is the events' own
    // cost, taken off each launch's figure in equal parts
    for (int t = 0; t <= 2 * trials && rc == FULLMAG_FDM_OK; ++t) {
        const bool inner = t <= trials;
        hipEventRecord(ev[0], h->stream);
        for (size_t i = 0; i < seq.size(); ++i) {
            with_args(h, [&](auto &A) { return fm::enqueue_kind(A, seq[i].first, seq[i].second, h->stream); });
            if (inner || i + 1 == seq.size()) hipEventRecord(ev[i + 1], h->stream);
        }
        if (hipEventSynchronize(ev.back()) != hipSuccess) {
            rc = fail(h, FULLMAG_FDM_ERR_CUDA, "timing sync");
            break;
        }
        if (!inner) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ev[0], ev.back());
            plain += 1000.0 * ms;
        } else if (t > 0) {
            float all = 0.f;
            hipEventElapsedTime(&all, ev[0], ev.back());
            with_events += 1000.0 * all;
            // the first launch of each kind in the trial step (stage 2; the
            // final x pass at the last stage)
            bool seen[7] = {false};
            for (size_t i = 0; i < seq.size(); ++i) {
                const int k = seq[i].first;
                if (seen[k]) continue;
                seen[k] = true;
                float ms = 0.f;
                hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
                sum[k] += 1000.0 * ms;
                ++cnt[k];
            }
        }
        *h->ctl_h = saved;
        if (!push_ctl(h)) rc = FULLMAG_FDM_ERR_CUDA;
    }
    for (auto &e : ev) hipEventDestroy(e);
    const double overhead = std::max(0.0, (with_events - plain) / trials / (double)seq.size());
    for (int k = 0; k < 7; ++k) out_us[k] = cnt[k] ? std::max(0.0, sum[k] / cnt[k] - overhead) : 0.0;
    if (n_out > 7) out_us[7] = plain / trials;
    if (n_out > 8) out_us[8] = overhead;
    *h->ctl_h = saved;
    if (!push_ctl(h) || !pull_ctl(h)) rc = FULLMAG_FDM_ERR_CUDA;
    h->obs_fresh = false;
    return rc;
}

}  // extern "C"


namespace {

// Times the x-pass launches of one trial step (stage-2 forward, joined,
// final inverse) for the current arguments; the state is not changed.
template <class T>
double time_x_passes(Backend *h, const fm::Args<T> &A, int reps, hipStream_t st) {
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    double total = 0.0;
    const int kinds[3] = {0, 6, 4};
    const double weight[3] = {1.0, (double)std::max(h->nstage - 2, 0), 1.0};
    for (int i = 0; i < 3; ++i) {
        if (fm::enqueue_kind(A, kinds[i], 2, st) != hipSuccess) continue;
        hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) fm::enqueue_kind(A, kinds[i], 2, st);
        hipEventRecord(e1, st);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        total += weight[i] * ms / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return total;
}

// Mean time of the 3-D y/z passes of one stage (forward y-lines, z pass,
// inverse y-lines), reps back-to-back launches of each on stream st.
template <class T>
double time_yz_passes(const fm::Args<T> &A, int reps, hipStream_t st) {
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    double total = 0.0;
    for (int kind = 1; kind <= 3; ++kind) {
        if (fm::enqueue_kind(A, kind, 2, st) != hipSuccess) continue;
        hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) fm::enqueue_kind(A, kind, 2, st);
        hipEventRecord(e1, st);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        total += ms / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return total;
}

// Chooses the 3-D y-line tile (threads per work-group: 128 or 256) and the
// z-pass tile (ky per work-group: the full, half or quarter count) by timing
// `apply`'d candidates with `time` (seconds); leaves the winner applied.
// Films and the Stockham y/z forms have no such tiles (false: nothing done).
template <class T, class Apply, class Time>
bool tune_yz(const fm::Args<T> &A0, Apply &&apply, Time &&time, int *yl_out, int *zf_out) {
    if (A0.nz == 1 || (!A0.ylfast && !A0.zfast)) return false;
    int zfull = 0;
    {
        int r1 = 0, r2 = 0;
        if (A0.zfast && fm::zfast_factors(A0.pz, &r1, &r2)) zfull = 256 / std::max(r1, r2);
    }
    double best = std::numeric_limits<double>::infinity();
    int best_yl = 0, best_zf = 0;
    for (int yl : {256, 128}) {
        if (!A0.ylfast && yl != 256) continue;
        for (int div : {1, 2, 4}) {
            const int zf = zfull ? zfull / div : 0;
            if (zfull && zf < 1) continue;
            if (!zfull && div != 1) continue;
            if (!apply(yl, zf)) return false;
            time(2);
            const double t = time(5);
            if (t < best) {
                best = t;
                best_yl = yl;
                best_zf = zf;
            }
        }
    }
    if (!apply(best_yl, best_zf)) return false;
    *yl_out = best_yl;
    *zf_out = best_zf;
    return true;
}

template <class T>
int autotune_typed(Backend *h, fm::Args<T> &A, int reps, int *out_rows, int *out_fs) {
    if (!A.demag) {
        if (out_rows) *out_rows = 0;
        if (out_fs) *out_fs = 0;
        return FULLMAG_FDM_OK;
    }
    if (!pull_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    const fm::Ctl saved = *h->ctl_h;
    HistoryGuard keep(h);
    if (!keep.ok()) return fail(h, FULLMAG_FDM_ERR_CUDA, "history copy allocation failed");
    const fm::Args<T> orig = A;
    double best = std::numeric_limits<double>::infinity();
    fm::Args<T> best_args = A;
    for (int fs = 1; fs >= 0; --fs) {
        for (int rows : {1, 2, 4}) {
            fm::Args<T> B = orig;
            const char *ce = fm::configure(B, h->lds_limit, rows, fs);
            if (ce && *ce) continue;
            if (fs == 1 && !B.xfast) continue;  // no four-step shape: the fs = 0 pass covers it
            if (!put_args(h, B, 0)) return FULLMAG_FDM_ERR_CUDA;
            time_x_passes(h, B, 2, h->stream);  // warm-up
            const double t = time_x_passes(h, B, reps, h->stream);
            if (t < best) {
                best = t;
                best_args = B;
            }
        }
    }
    A = best_args;
    const char *ce = fm::configure(A, h->lds_limit, A.rows_per_wg, A.xfast ? 1 : 0);  // kernel LDS attributes
    if (ce && *ce) return fail(h, FULLMAG_FDM_ERR_INTERNAL, ce);
    if (!put_args(h, A, 0)) return FULLMAG_FDM_ERR_CUDA;
    // 3-D: the y-line and z-pass tiles, timed on the y/z passes
    int yl = A.yl_threads, zf = A.zf_ub;
    {
        auto apply = [&](int y, int z) {
            A.yl_threads = y;
            A.zf_ub = z;
            return put_args(h, A, 0);
        };
        auto time = [&](int r) { return time_yz_passes(A, r, h->stream); };
        if (!tune_yz(A, apply, time, &yl, &zf) && !h->err.empty()) return FULLMAG_FDM_ERR_CUDA;
        A.yl_threads = yl;
        A.zf_ub = zf;
        if (!put_args(h, A, 0)) return FULLMAG_FDM_ERR_CUDA;
    }
    // graphs were captured with the old launch geometry
    if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
    if (h->graph_chunk) hipGraphExecDestroy(h->graph_chunk);
    if (h->graph) hipGraphDestroy(h->graph);
    h->graph_exec = h->graph_chunk = nullptr;
    h->graph = nullptr;
    h->graph_tried = false;
    *h->ctl_h = saved;
    if (!push_ctl(h) || !eval_current(h, false)) return FULLMAG_FDM_ERR_CUDA;
    tuned_store(h->fp64, (int)h->nx, (int)h->ny, (int)h->nz, A.method, tune_feat(A), A.rows_per_wg, A.xfast ? 1 : 0,
                A.yl_threads, A.zf_ub);
    if (out_rows) *out_rows = A.rows_per_wg;
    if (out_fs) *out_fs = A.xfast ? 1 : 0;
    return FULLMAG_FDM_OK;
}

}  // namespace

extern "C" int fullmag_fdm_backend_autotune(fullmag_fdm_backend *h, int reps, int *out_rows_per_group,
                                            int *out_four_step_x) {
    if (!h || reps < 1) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    h->err.clear();
    hipSetDevice(h->device);
    return with_args(h, [&](auto &A) { return autotune_typed(h, A, reps, out_rows_per_group, out_four_step_x); });
}

// ------------------------------------------------------------ adjoint --
namespace {

// Captures the launches `f` enqueues on the handle's stream into an
// executable graph (nullptr on failure).
template <class F>
hipGraphExec_t capture_graph(Backend *h, F &&f) {
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    const bool ok = f();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    hipGraphExec_t ex = nullptr;
    if (ok && e == hipSuccess && g && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
    if (g) hipGraphDestroy(g);
    (void)hipGetLastError();
    return ex;
}

void adj_free(Backend *h) {
    for (void *p : h->adj.bufs) hipFree(p);
    h->adj = Backend::Adj{};
}

void *adj_alloc(Backend *h, size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    h->adj.bufs.push_back(p);
    return p;
}

template <class T>
fm::AdjArgs<T> adj_args(Backend *h) {
    auto &D = h->adj;
    fm::AdjArgs<T> J{};
    const int ns = h->nstage;
    J.rec_y0 = static_cast<const T *>(D.rec_y);
    J.rec_h0 = static_cast<const T *>(D.rec_h);
    J.rec_vn0 = static_cast<const T *>(D.rec_vn);
    J.step = D.step;
    J.seg = (long long)D.rec_seg;
    J.sv = (long long)(ns - 1) * 3 * (long long)h->n;
    J.ss = (long long)(ns - 1) * (long long)h->n;
    J.tn = D.tn;
    J.first = (long long)D.first;
    J.cot_m = static_cast<double *>(D.cot_m);
    J.cot_k = static_cast<double *>(D.cot_k);
    J.beta = static_cast<double *>(D.beta);
    J.grad_ms = D.grad_ms;
    J.grad_amp = D.grad_amp;
    J.obj_part = D.obj_part;
    J.obj_w = static_cast<const T *>(D.obj_w);
    J.obj_comp = (int)D.comp;
    J.obj_scale = D.obj_scale;
    J.dt = D.dt;
    return J;
}

// Recording form of the arguments: the trial step's kernels store the
// stage states, fields and pre-normalisation norms in record slot
// (device step index mod seg).
template <class T>
fm::Args<T> adj_record_args(Backend *h, const fm::Args<T> &A) {
    const auto &D = h->adj;
    const int ns = h->nstage;
    fm::Args<T> B = A;
    B.rec_y = static_cast<T *>(D.rec_y);
    B.rec_h = static_cast<T *>(D.rec_h);
    B.rec_vn = static_cast<T *>(D.rec_vn);
    B.rec_step = D.step;
    B.rec_seg = (long long)D.rec_seg;
    B.rec_sv = (long long)(ns - 1) * 3 * (long long)h->n;
    B.rec_ss = (long long)(ns - 1) * (long long)h->n;
    return B;
}

// FP64 demag context of an FP32 system (same grid, FP64 tensor, spectrum
// and twiddles): the adjoint's field cotangent beta is convolved there.  The
// cotangent's far field from a small source (a pulse region acting on
// distant probe cells) is a small difference of FFT terms of the size of its
// near field; in FP32 it keeps ~1e-7 of that near field, in FP64 ~1e-16.
bool build_f64_context(Backend *h) {
    if (h->f64_ready) return true;
    const fm::Args<float> &A = h->af;
    if (h->fp64 || !A.demag || A.lmix || A.gtens || h->red64.empty()) return false;
    fm::Args<double> F{};
    F.nx = A.nx;
    F.ny = A.ny;
    F.nz = A.nz;
    F.ncell = A.ncell;
    F.nrows = A.nrows;
    F.nkx = A.nkx;
    F.ly = A.ly;
    F.nzs = A.nzs;
    F.px = A.px;
    F.py = A.py;
    F.pz = A.pz;
    F.inv_p = A.inv_p;
    F.fx = A.fx;
    F.fy = A.fy;
    F.fz = A.fz;
    F.demag = 1;
    F.method = A.method;
    F.ctl = A.ctl;
    const int nx = (int)h->nx, ny = (int)h->ny, nz = (int)h->nz;
    const std::vector<double> &red = h->red64;
    std::vector<double> rt(red.size());
    if (nz == 1) {
        rt = red;
    } else {
        const size_t X = nx + 1, Y = ny + 1, Z = nz + 1;
        for (size_t c = 0; c < 6; ++c)
            for (size_t kx = 0; kx < X; ++kx)
                for (size_t ky = 0; ky < Y; ++ky)
                    for (size_t kz = 0; kz < Z; ++kz)
                        rt[((kx * Z + kz) * Y + ky) * 6 + c] = red[((c * X + kx) * Y + ky) * Z + kz];
    }
    double *nt = static_cast<double *>(dalloc(h, rt.size() * sizeof(double)));
    if (!nt || !hip_ok(h, hipMemcpy(nt, rt.data(), rt.size() * sizeof(double), hipMemcpyHostToDevice),
                       "FP64 demag context"))
        return false;
    F.ntens = nt;
    F.spec = dalloc(h, 3ull * F.nzs * F.nkx * F.ly * 2 * sizeof(double));
    F.twx = upload_twiddles<double>(h, F.px);
    F.twy = upload_twiddles<double>(h, F.py);
    F.twz = F.pz > 1 ? upload_twiddles<double>(h, F.pz) : F.twy;
    if (!F.spec || !F.twx || !F.twy || !F.twz) return hip_ok(h, hipErrorOutOfMemory, "FP64 demag context");
    if (!four_step_twiddles<double>(h, F.py, &F.tw4y) ||
        (nz > 1 && !four_step_twiddles<double>(h, F.py, &F.tw4y_f, false, true)) ||
        !four_step_twiddles<double>(h, F.px, &F.tw4x) ||
        (nz > 1 && !four_step_twiddles<double>(h, F.pz, &F.tw4z, true)))
        return false;
    const char *ce = fm::configure(F, h->lds_limit);
    if (ce && *ce) return (h->err = std::string("FP64 demag context: ") + ce, false);
    void *ad = dalloc(h, sizeof(fm::Args<double>));
    if (!ad) return hip_ok(h, hipErrorOutOfMemory, "FP64 demag context");
    F.self_d = static_cast<const fm::Args<double> *>(ad);
    F.nsys = 1;
    if (!hip_ok(h, hipMemcpy(ad, &F, sizeof(F), hipMemcpyHostToDevice), "FP64 demag context") ||
        !hip_ok(h, fm::adj_configure_f64(F), "FP64 demag context LDS limits"))
        return false;
    h->f64 = F;
    h->f64_ready = true;
    return true;
}

// Reverse sweep of stage j in the system's precision or, for an FP32 system
// with beta64, with the field cotangent convolved in the FP64 context.
template <class T>
hipError_t adj_stage(Backend *h, const fm::Args<T> &A, const fm::AdjArgs<T> &J, int j) {
    if constexpr (sizeof(T) == 4) {
        if (h->adj.beta64) return fm::enqueue_adj_stage_f64(A, h->f64, J, j, h->stream);
    }
    return fm::enqueue_adj_stage(A, J, j, h->stream);
}

template <class T>
int adj_forward_typed(Backend *h, fm::Args<T> &A, const fullmag_fdm_adjoint_desc *d, double *out_j) {
    adj_free(h);
    auto &D = h->adj;
    const uint64_t n = h->n, steps = d->steps;
    const int ns = h->nstage;
    D.steps = steps;
    D.dt = d->dt;
    D.first = d->objective_first_step;
    D.comp = d->objective_component;
    D.obj_scale = 1.0 / (double)(steps - D.first + 1);
    D.seg = d->checkpoint_interval ? d->checkpoint_interval : (uint64_t)std::ceil(std::sqrt((double)steps));
    D.seg = std::min<uint64_t>(std::max<uint64_t>(D.seg, 1), steps);
    D.ncp = (steps + D.seg - 1) / D.seg;
    // trailing segments recorded by the forward itself (no recompute in the
    // backward): as many as the record budget holds, at least the last one
    {
        const uint64_t per_step = (uint64_t)(ns - 1) * (2 * 3 * n + n) * sizeof(T);
        uint64_t budget = 0;
        const bool has_budget = d->struct_size >= offsetof(fullmag_fdm_adjoint_desc, record_budget_bytes) +
                                                      sizeof(d->record_budget_bytes);
        if (has_budget && d->record_budget_bytes) {
            budget = d->record_budget_bytes;
        } else {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) == hipSuccess) budget = fr / 2;
        }
        const uint64_t seg_bytes = std::max<uint64_t>(per_step * D.seg, 1);
        D.krec = std::min<uint64_t>(std::max<uint64_t>(budget / seg_bytes, 1), D.ncp);
        D.rec_seg = D.krec * D.seg + 1;
    }
    D.blocks = fm::adj_blocks(A);
    // FP32 systems convolve the field cotangent in FP64 unless the caller
    // asks for the state precision (or no FP64 context can be built)
    D.beta64 = false;
    if constexpr (sizeof(T) == 4) {
        if (A.demag && !(d->flags & FULLMAG_FDM_ADJOINT_BETA_STATE_PRECISION)) {
            if (!h->f64_ready) {
                const std::string keep = h->err;
                if (!build_f64_context(h) && h->err != keep) return FULLMAG_FDM_ERR_CUDA;
            }
            D.beta64 = h->f64_ready;
            if (D.beta64) D.blocks = std::max<long long>(D.blocks, fm::num_partials(h->f64));
        }
    }
    const size_t v3 = 3 * n * sizeof(T), slots = (size_t)D.rec_seg * (ns - 1);
    D.cp = adj_alloc(h, D.ncp * v3);
    D.rec_y = adj_alloc(h, slots * v3);
    D.rec_h = adj_alloc(h, slots * v3);
    D.rec_vn = adj_alloc(h, slots * n * sizeof(T));
    D.cot_m = adj_alloc(h, 3 * n * sizeof(double));
    D.cot_k = adj_alloc(h, fm::kMaxStages * 3 * n * sizeof(double));
    D.beta = adj_alloc(h, 3 * n * sizeof(double));
    D.obj_w = adj_alloc(h, n * sizeof(T));
    D.m_end = adj_alloc(h, v3);
    D.grad_ms = static_cast<double *>(adj_alloc(h, n * sizeof(double)));
    D.grad_amp = static_cast<double *>(adj_alloc(h, (size_t)D.blocks * 3 * sizeof(double)));
    D.obj_part = static_cast<double *>(adj_alloc(h, (size_t)D.blocks * 3 * sizeof(double)));
    D.tn = static_cast<double *>(adj_alloc(h, (steps + 1) * sizeof(double)));
    D.step = static_cast<long long *>(adj_alloc(h, sizeof(long long)));
    if (!D.cp || !D.rec_y || !D.rec_h || !D.rec_vn || !D.cot_m || !D.cot_k || !D.beta || !D.obj_w || !D.m_end ||
        !D.grad_ms || !D.grad_amp || !D.obj_part || !D.tn || !D.step) {
        adj_free(h);
        return fail(h, FULLMAG_FDM_ERR_CUDA, "adjoint: device allocation failed");
    }
    std::vector<T> w(n);
    for (uint64_t i = 0; i < n; ++i) w[i] = (T)d->objective_weight[i];
    if (!hip_ok(h, hipMemcpy(D.obj_w, w.data(), n * sizeof(T), hipMemcpyHostToDevice), "adjoint weights") ||
        !hip_ok(h, fm::adj_configure(A), "adjoint LDS limits") || !pull_ctl(h))
        return FULLMAG_FDM_ERR_CUDA;
    fm::Ctl &c = *h->ctl_h;
    D.t0 = c.t;
    D.parity0 = c.parity;
    const size_t n3 = 3 * n;
    if (!hip_ok(h, hipMemcpyAsync(D.cp, A.m_all + (size_t)c.parity * n3, v3, hipMemcpyDeviceToDevice, h->stream),
                "adjoint checkpoint") ||
        !hip_ok(h, hipMemsetAsync(D.obj_part, 0, (size_t)D.blocks * 3 * sizeof(double), h->stream),
                "adjoint objective clear"))
        return FULLMAG_FDM_ERR_CUDA;
    c.loop = 1;
    c.t_end = std::numeric_limits<double>::infinity();
    c.step_end = c.step + steps;
    c.dt_carry = D.dt;
    c.dt = D.dt;
    c.done = 0;
    // host replica of the device time sequence t_n = t_{n-1} + dt (the
    // controller adds dt in FP64 after every accepted step)
    std::vector<double> tn(steps + 1);
    tn[0] = D.t0;
    for (uint64_t s = 1; s <= steps; ++s) tn[s] = tn[s - 1] + D.dt;
    if (!hip_ok(h, hipMemcpyAsync(D.tn, tn.data(), tn.size() * sizeof(double), hipMemcpyHostToDevice, h->stream),
                "adjoint step times"))
        return FULLMAG_FDM_ERR_CUDA;
    fm::AdjArgs<T> J = adj_args<T>(h);
    J.obj_from = (long long)(c.step + D.first);
    if (!push_ctl(h) || !hip_ok(h, hipStreamSynchronize(h->stream), "adjoint setup")) return FULLMAG_FDM_ERR_CUDA;
    h->obs_fresh = false;
    auto one_step = [&]() {
        return hip_ok(h, fm::enqueue_trial(A, ns, h->stream), "adjoint forward step") &&
               hip_ok(h, fm::enqueue_adj_objective(A, J, h->stream), "adjoint objective");
    };
    // the last krec segments run in recording form (the arguments of the
    // backward's recompute, record ring of krec*seg steps), so that the
    // backward reverses them from this record instead of recomputing them;
    // when the budget holds every segment nothing is recomputed
    fm::Args<T> B = adj_record_args(h, A);
    if (!put_args(h, B, 2)) return FULLMAG_FDM_ERR_CUDA;
    auto rec_step = [&]() {
        return hip_ok(h, fm::enqueue_trial(B, ns, h->stream), "adjoint forward step (recording)") &&
               hip_ok(h, fm::enqueue_adj_objective(A, J, h->stream), "adjoint objective");
    };
    const uint64_t n_last = (D.ncp - D.krec) * D.seg;
    const size_t sv_bytes = (size_t)(ns - 1) * v3;  // one step slot of rec_y
    hipGraphExec_t g = capture_graph(h, one_step);
    hipGraphExec_t g_rec = capture_graph(h, rec_step);
    int par = D.parity0;
    bool ok = true;
    for (uint64_t s = 1; s <= steps && ok; ++s) {
        if (s - 1 == n_last) {
            // entering the recorded segments at state n_last: the record
            // starts with that state and its k1, as the backward's restart does
            char *slot0 = static_cast<char *>(D.rec_y) + (size_t)(n_last % D.rec_seg) * sv_bytes;
            ok = hip_ok(h, fm::enqueue_adj_step(D.step, (long long)n_last, 0, h->stream), "adjoint step") &&
                 hip_ok(h, hipMemcpyAsync(slot0, A.m_all + (size_t)par * n3, v3, hipMemcpyDeviceToDevice, h->stream),
                        "adjoint record start") &&
                 hip_ok(h, fm::enqueue_eval(B, 1, 0, h->stream), "adjoint forward (k1 record)");
            if (!ok) break;
        }
        if (s - 1 >= n_last) ok = g_rec ? hip_ok(h, hipGraphLaunch(g_rec, h->stream), "adjoint forward graph") : rec_step();
        else ok = g ? hip_ok(h, hipGraphLaunch(g, h->stream), "adjoint forward graph") : one_step();
        par ^= 1;
        if (ok && s % D.seg == 0 && s < steps)
            ok = hip_ok(h,
                        hipMemcpyAsync(static_cast<char *>(D.cp) + (s / D.seg) * v3, A.m_all + (size_t)par * n3, v3,
                                       hipMemcpyDeviceToDevice, h->stream),
                        "adjoint checkpoint");
    }
    if (g) hipGraphExecDestroy(g);
    if (g_rec) hipGraphExecDestroy(g_rec);
    if (!ok) return FULLMAG_FDM_ERR_CUDA;
    std::vector<double> part((size_t)D.blocks * 3);
    if (!hip_ok(h, hipMemcpyAsync(part.data(), D.obj_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  h->stream),
                "adjoint objective readback") ||
        !pull_ctl(h))
        return FULLMAG_FDM_ERR_CUDA;
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(h, FULLMAG_FDM_ERR_CUDA, std::string("device error: ") + hipGetErrorString(le));
    double jsum = 0.0;
    for (long long b = 0; b < D.blocks; ++b) jsum += part[(size_t)b * 3];
    c.loop = 0;
    c.done = 0;
    if (!push_ctl(h) || !pull_ctl(h)) return FULLMAG_FDM_ERR_CUDA;
    if (!hip_ok(h, hipMemcpyAsync(D.m_end, A.m_all + (size_t)c.parity * n3, v3, hipMemcpyDeviceToDevice, h->stream),
                "adjoint final state"))
        return FULLMAG_FDM_ERR_CUDA;
    D.ctl_end = c;
    D.last_rec = true;
    D.ready = true;
    if (out_j) *out_j = jsum * D.obj_scale;
    return FULLMAG_FDM_OK;
}

template <class T>
int adj_backward_typed(Backend *h, fm::Args<T> &A, double *grad_ms, double *grad_amp) {
    auto &D = h->adj;
    const uint64_t n = h->n;
    const int ns = h->nstage;
    const size_t n3 = 3 * n, v3 = n3 * sizeof(T);
    if (!hip_ok(h, hipMemsetAsync(D.grad_ms, 0, n * sizeof(double), h->stream), "adjoint clear") ||
        !hip_ok(h, hipMemsetAsync(D.grad_amp, 0, (size_t)D.blocks * 3 * sizeof(double), h->stream), "adjoint clear") ||
        !hip_ok(h, hipMemsetAsync(D.cot_m, 0, 3 * n * sizeof(double), h->stream), "adjoint clear"))
        return FULLMAG_FDM_ERR_CUDA;
    const fm::AdjArgs<T> J = adj_args<T>(h);
    fm::Args<T> B = adj_record_args(h, A);
    if (!put_args(h, B, 2)) return FULLMAG_FDM_ERR_CUDA;
    auto rec_step = [&]() { return hip_ok(h, fm::enqueue_trial(B, ns, h->stream), "adjoint recompute"); };
    auto rev_step = [&]() {
        if (!hip_ok(h, fm::enqueue_adj_init(A, J, ns, h->stream), "adjoint init")) return false;
        for (int j = ns - 1; j >= 1; --j)
            if (!hip_ok(h, adj_stage(h, A, J, j), "adjoint stage")) return false;
        return hip_ok(h, fm::enqueue_adj_step(D.step, -1, 1, h->stream), "adjoint step");
    };
    hipGraphExec_t g_rec = capture_graph(h, rec_step);
    hipGraphExec_t g_rev = capture_graph(h, rev_step);
    auto run = [&](hipGraphExec_t g, auto &f) {
        return g ? hip_ok(h, hipGraphLaunch(g, h->stream), "adjoint graph") : f();
    };
    fm::Ctl &c = *h->ctl_h;
    h->obs_fresh = false;
    bool ok = true;
    for (uint64_t k = D.ncp; k-- > 0 && ok;) {
        const uint64_t n0 = k * D.seg, n1 = std::min<uint64_t>(n0 + D.seg, D.steps);
        // the pinned control block is read by the queued upload only when the
        // stream reaches it: drain the stream before rewriting it
        if (!hip_ok(h, hipStreamSynchronize(h->stream), "adjoint sync")) {
            ok = false;
            break;
        }
        if (k + D.krec >= D.ncp && D.last_rec) {
            // recorded by the forward run: reverse it directly
            ok = hip_ok(h, hipMemsetAsync(&h->ctl_d->done, 0, sizeof(int), h->stream), "adjoint loop flag") &&
                 hip_ok(h, fm::enqueue_adj_step(D.step, (long long)n1 - 1, 0, h->stream), "adjoint step");
            for (uint64_t i = n0; i < n1 && ok; ++i) ok = run(g_rev, rev_step);
            continue;
        }
        // recompute the segment from its checkpoint, recording every stage
        c = D.ctl_end;
        c.t = D.t0;
        for (uint64_t s = 0; s < n0; ++s) c.t += D.dt;
        c.parity = 0;
        c.loop = 1;
        c.t_end = std::numeric_limits<double>::infinity();
        c.step_end = c.step + (n1 - n0);
        c.dt_carry = D.dt;
        c.dt = D.dt;
        c.done = 0;
        const char *cpk = static_cast<const char *>(D.cp) + k * v3;
        char *slot0 = static_cast<char *>(D.rec_y) + (size_t)(n0 % D.rec_seg) * (size_t)(ns - 1) * v3;
        ok = hip_ok(h, hipMemcpyAsync(A.m_all, cpk, v3, hipMemcpyDeviceToDevice, h->stream), "adjoint restart") &&
             hip_ok(h, hipMemcpyAsync(slot0, cpk, v3, hipMemcpyDeviceToDevice, h->stream), "adjoint restart") &&
             push_ctl(h) && hip_ok(h, fm::enqueue_adj_step(D.step, (long long)n0, 0, h->stream), "adjoint step") &&
             hip_ok(h, fm::enqueue_eval(B, 1, 0, h->stream), "adjoint recompute (k1)");
        for (uint64_t i = n0; i < n1 && ok; ++i) ok = run(g_rec, rec_step);
        if (!ok) break;
        // the recompute ends with the device loop marked done; the y/z passes
        // that the reverse sweep shares with the forward step return early
        // while it is set
        ok = hip_ok(h, hipMemsetAsync(&h->ctl_d->done, 0, sizeof(int), h->stream), "adjoint loop flag") &&
             hip_ok(h, fm::enqueue_adj_step(D.step, (long long)n1 - 1, 0, h->stream), "adjoint step");
        for (uint64_t i = n0; i < n1 && ok; ++i) ok = run(g_rev, rev_step);
    }
    if (g_rec) hipGraphExecDestroy(g_rec);
    if (g_rev) hipGraphExecDestroy(g_rev);
    if (!ok) return FULLMAG_FDM_ERR_CUDA;
    std::vector<double> gms(n), gamp((size_t)D.blocks * 3);
    if (!hip_ok(h, hipMemcpyAsync(gms.data(), D.grad_ms, n * sizeof(double), hipMemcpyDeviceToHost, h->stream),
                "gradient readback") ||
        !hip_ok(h, hipMemcpyAsync(gamp.data(), D.grad_amp, gamp.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  h->stream),
                "gradient readback") ||
        !hip_ok(h, hipStreamSynchronize(h->stream), "adjoint sync"))
        return FULLMAG_FDM_ERR_CUDA;
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(h, FULLMAG_FDM_ERR_CUDA, std::string("device error: ") + hipGetErrorString(le));
    // restore the final state of the forward run and its k1
    c = D.ctl_end;
    if (!hip_ok(h, hipMemcpyAsync(A.m_all + (size_t)c.parity * n3, D.m_end, v3, hipMemcpyDeviceToDevice, h->stream),
                "adjoint final state") ||
        !push_ctl(h) || !eval_current(h, false))
        return FULLMAG_FDM_ERR_CUDA;
    // the recomputed segments reused the record slots of the recorded ones
    if (D.ncp > D.krec) D.last_rec = false;
    if (grad_ms)
        for (uint64_t i = 0; i < n; ++i) grad_ms[i] = gms[i];
    if (grad_amp) {
        grad_amp[0] = grad_amp[1] = grad_amp[2] = 0.0;
        for (long long b = 0; b < D.blocks; ++b)
            for (int q = 0; q < 3; ++q) grad_amp[q] += gamp[(size_t)b * 3 + q];
    }
    return FULLMAG_FDM_OK;
}

}  // namespace

extern "C" {

int fullmag_fdm_adjoint_forward(fullmag_fdm_backend *h, const fullmag_fdm_adjoint_desc *d, double *out_objective) {
    if (!h || !d) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (d->struct_size < offsetof(fullmag_fdm_adjoint_desc, record_budget_bytes))
        return fail(h, FULLMAG_FDM_ERR_INVALID, "fullmag_fdm_adjoint_desc.struct_size too small");
    if (h->adaptive || h->multistep)
        return fail(h, FULLMAG_FDM_ERR_INVALID, "adjoint needs a one-step fixed-step integrator (rk4 or heun)");
    if (h->fp64 ? h->ad.xt : h->af.xt)
        return fail(h, FULLMAG_FDM_ERR_INVALID,
                    "adjoint covers exchange, demag, uniaxial anisotropy and Zeeman terms only");
    if (d->steps == 0 || !(d->dt > 0.0) || !std::isfinite(d->dt))
        return fail(h, FULLMAG_FDM_ERR_INVALID, "adjoint: steps must be >= 1 and dt positive and finite");
    if (d->objective_first_step < 1 || d->objective_first_step > d->steps)
        return fail(h, FULLMAG_FDM_ERR_INVALID, "adjoint: objective_first_step must be in 1..steps");
    if (d->objective_component > 2) return fail(h, FULLMAG_FDM_ERR_INVALID, "adjoint: objective_component must be 0..2");
    if (!d->objective_weight || d->objective_weight_len != h->n)
        return fail(h, FULLMAG_FDM_ERR_INVALID, "adjoint: objective_weight_len must equal the cell count");
    const bool lut = h->fp64 ? h->ad.region != nullptr : h->af.region != nullptr;
    if (lut) return fail(h, FULLMAG_FDM_ERR_INVALID, "adjoint: region exchange tables are not supported");
    h->err.clear();
    hipSetDevice(h->device);
    h->adj.ready = false;
    return with_args(h, [&](auto &A) { return adj_forward_typed(h, A, d, out_objective); });
}

int fullmag_fdm_adjoint_backward(fullmag_fdm_backend *h, double *grad_ms, double *grad_field_amplitude) {
    if (!h) return FULLMAG_FDM_ERR_INVALID;
    if (!h->ok) return fail(h, FULLMAG_FDM_ERR_INVALID, "backend not initialised");
    if (!h->adj.ready) return fail(h, FULLMAG_FDM_ERR_INVALID, "adjoint_backward needs a preceding adjoint_forward");
    h->err.clear();
    hipSetDevice(h->device);
    return with_args(h, [&](auto &A) { return adj_backward_typed(h, A, grad_ms, grad_field_amplitude); });
}

}  // extern "C"

// ------------------------------------------------------------ batches --
struct fullmag_fdm_batch {
    std::string err;
    std::vector<fullmag_fdm_backend *> sys;
    struct Group {
        std::vector<size_t> members;
        bool fp64 = false;
        int nstage = 0;
        void *args_d = nullptr;  // members' kernel arguments, one per system
        hipStream_t stream = nullptr;
        hipGraphExec_t graph = nullptr;
        fm::Args<float> gf{};
        fm::Args<double> gd{};
        bool finished = false;
        // two polling chunks in flight: each ends with the members' control
        // blocks read back into its flag slot and an event
        hipEvent_t ev[2] = {nullptr, nullptr};
        int next = 0;  // slot of the older chunk in flight
    };
    std::vector<Group> groups;
    fm::Ctl *flags = nullptr;  // pinned, [2][systems]: control blocks read back per chunk slot
};

namespace {

using Batch = fullmag_fdm_batch;

// Launch geometry that must agree for systems to share launches.
template <class T>
std::vector<long long> group_key(const fm::Args<T> &A, bool fp64) {
    // the y/z stage of a launch is chosen from the group's first member:
    // multilayer mixing, periodic (complex) spectra and the z-neighbour rule
    // must agree as well
    return {fp64, A.nx, A.ny, A.nz, A.demag, A.method, A.region != nullptr, A.rows_per_wg, A.xfast, A.yfast,
            A.ylfast, A.zfast, A.ty, A.nseq_x, A.xt, A.has_th, A.lmix != nullptr, A.layered, A.pbc,
            A.gtens != nullptr, A.xpf, A.yl_threads, A.zf_ub};
}

template <class T>
fm::Args<T> &group_args(Batch::Group &g);
template <>
fm::Args<float> &group_args<float>(Batch::Group &g) { return g.gf; }
template <>
fm::Args<double> &group_args<double>(Batch::Group &g) { return g.gd; }

// Uploads the members' current kernel arguments into the group array and
// captures one polling chunk of trial steps of the whole group (once).
template <class T>
bool group_prepare(Batch *b, Batch::Group &g, bool capture = true) {
    const size_t ns = g.members.size();
    if (!g.args_d && hipMalloc(&g.args_d, ns * sizeof(fm::Args<T>)) != hipSuccess)
        return (b->err = "batch: argument allocation failed", false);
    std::vector<fm::Args<T>> host(ns);
    for (size_t i = 0; i < ns; ++i) {
        Backend *h = b->sys[g.members[i]];
        host[i] = *reinterpret_cast<fm::Args<T> *>(h->fp64 ? (void *)&h->ad : (void *)&h->af);
    }
    if (hipStreamSynchronize(g.stream) != hipSuccess ||
        hipMemcpy(g.args_d, host.data(), ns * sizeof(fm::Args<T>), hipMemcpyHostToDevice) != hipSuccess)
        return (b->err = "batch: argument upload failed", false);
    fm::Args<T> &G = group_args<T>(g);
    G = host[0];
    G.self_d = static_cast<const fm::Args<T> *>(g.args_d);
    G.nsys = (int)ns;
    if (capture && !g.graph) {
        if (hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            hipError_t e = hipSuccess;
            for (int a = 0; a < kAttemptChunk && e == hipSuccess; ++a) e = fm::enqueue_trial(G, g.nstage, g.stream);
            hipGraph_t gr = nullptr;
            const hipError_t ec = hipStreamEndCapture(g.stream, &gr);
            if (e == hipSuccess && ec == hipSuccess && gr &&
                hipGraphInstantiate(&g.graph, gr, nullptr, nullptr, 0) != hipSuccess)
                g.graph = nullptr;
            if (gr) hipGraphDestroy(gr);
            (void)hipGetLastError();
        }
    }
    return true;
}

// One polling chunk (kAttemptChunk trial steps) of a group.
// Chooses the x tile (rows per work-group, FFT form) of a group by timing
// the group's batched x-pass launches for each candidate; every member gets
// the winner, which is also recorded for later creates of that shape.
// Members' states are not changed.
template <class T>
bool group_tune(Batch *b, Batch::Group &g) {
    Backend *h0 = b->sys[g.members[0]];
    fm::Args<T> &A0 = *reinterpret_cast<fm::Args<T> *>(h0->fp64 ? (void *)&h0->ad : (void *)&h0->af);
    if (!A0.demag) return true;
    auto apply = [&](int rows, int fs) {
        for (size_t m : g.members) {
            Backend *h = b->sys[m];
            fm::Args<T> &A = *reinterpret_cast<fm::Args<T> *>(h->fp64 ? (void *)&h->ad : (void *)&h->af);
            const char *ce = fm::configure(A, h->lds_limit, rows, fs);
            if ((ce && *ce) || !put_args(h, A, 0)) return false;
        }
        return true;
    };
    double best = std::numeric_limits<double>::infinity();
    int best_rows = A0.rows_per_wg, best_fs = A0.xfast ? 1 : 0;
    for (int fs = 1; fs >= 0; --fs)
        for (int rows : {1, 2, 4}) {
            if (!apply(rows, fs)) continue;
            if (fs == 1 && !A0.xfast) continue;
            if (!group_prepare<T>(b, g, false)) return false;
            time_x_passes(h0, group_args<T>(g), 2, g.stream);
            const double t = time_x_passes(h0, group_args<T>(g), 5, g.stream);
            if (t < best) {
                best = t;
                best_rows = rows;
                best_fs = fs;
            }
        }
    if (!apply(best_rows, best_fs)) return (b->err = "batch: tile configuration failed", false);
    // 3-D: the y-line and z-pass tiles, timed on the group's y/z launches
    int yl = A0.yl_threads, zf = A0.zf_ub;
    {
        auto apply_yz = [&](int y, int z) {
            for (size_t m : g.members) {
                Backend *h = b->sys[m];
                fm::Args<T> &A = *reinterpret_cast<fm::Args<T> *>(h->fp64 ? (void *)&h->ad : (void *)&h->af);
                A.yl_threads = y;
                A.zf_ub = z;
                if (!put_args(h, A, 0)) return false;
            }
            return group_prepare<T>(b, g, false);
        };
        auto time = [&](int r) { return time_yz_passes(group_args<T>(g), r, g.stream); };
        tune_yz(A0, apply_yz, time, &yl, &zf);
        if (!apply_yz(yl, zf)) return (b->err = "batch: tile configuration failed", false);
    }
    // k1 of the initial states with the chosen tiles (as a create with them gives)
    for (size_t m : g.members)
        if (!eval_current(b->sys[m], false)) return (b->err = "batch: evaluation failed", false);
    tuned_store(h0->fp64, (int)h0->nx, (int)h0->ny, (int)h0->nz, A0.method, tune_feat(A0), best_rows, best_fs, yl, zf);
    return true;
}

// One polling chunk of a group into flag slot `slot`: kAttemptChunk trial
// steps, the members' control blocks read back, the slot's event.
template <class T>
bool group_launch(Batch *b, Batch::Group &g, int slot) {
    if (g.graph) {
        if (hipGraphLaunch(g.graph, g.stream) != hipSuccess) return (b->err = "batch: graph launch", false);
    } else {
        for (int a = 0; a < kAttemptChunk; ++a)
            if (fm::enqueue_trial(group_args<T>(g), g.nstage, g.stream) != hipSuccess)
                return (b->err = "batch: launch failed", false);
    }
    fm::Ctl *f = b->flags + (size_t)slot * b->sys.size();
    for (size_t m : g.members)
        if (hipMemcpyAsync(&f[m], b->sys[m]->ctl_d, sizeof(fm::Ctl), hipMemcpyDeviceToHost, g.stream) != hipSuccess)
            return (b->err = "batch: control readback failed", false);
    if (hipEventRecord(g.ev[slot], g.stream) != hipSuccess) return (b->err = "batch: event record failed", false);
    return true;
}

}  // namespace

extern "C" {

fullmag_fdm_batch *fullmag_fdm_batch_create(const fullmag_fdm_plan_desc *const *plans,
                                            const fullmag_fdm_plan_ext *const *exts, uint64_t count) {
    if (!plans || count == 0) return nullptr;
    Batch *b = new (std::nothrow) Batch();
    if (!b) return nullptr;
    for (uint64_t i = 0; i < count; ++i) {
        Backend *h = fullmag_fdm_backend_create_ext(plans[i], exts ? exts[i] : nullptr);
        if (!h) {
            b->err = "batch: system " + std::to_string(i) + ": create returned NULL";
            return b;
        }
        b->sys.push_back(h);
        if (!h->ok) {
            b->err = "batch: system " + std::to_string(i) + ": " + h->err;
            return b;
        }
    }
    std::vector<std::vector<long long>> keys;
    for (size_t i = 0; i < b->sys.size(); ++i) {
        Backend *h = b->sys[i];
        const auto key = h->fp64 ? group_key(h->ad, true) : group_key(h->af, false);
        size_t gi = 0;
        while (gi < keys.size() && keys[gi] != key) ++gi;
        if (gi == keys.size()) {
            keys.push_back(key);
            b->groups.emplace_back();
            b->groups.back().fp64 = h->fp64;
            b->groups.back().nstage = h->nstage;
        }
        b->groups[gi].members.push_back(i);
    }
    hipSetDevice(b->sys[0]->device);
    for (auto &g : b->groups) {
        if (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&g.ev[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&g.ev[1], hipEventDisableTiming) != hipSuccess) {
            b->err = "batch: stream creation failed";
            return b;
        }
        // x tile of the group, timed on the group's own batched launches
        const bool ok = g.fp64 ? group_tune<double>(b, g) : group_tune<float>(b, g);
        if (!ok) return b;
    }
    if (hipHostMalloc((void **)&b->flags, 2 * b->sys.size() * sizeof(fm::Ctl), hipHostMallocDefault) != hipSuccess)
        b->err = "batch: pinned allocation failed";
    return b;
}

int fullmag_fdm_batch_groups(const fullmag_fdm_batch *b, uint64_t *out) {
    if (!b || !out) return FULLMAG_FDM_ERR_INVALID;
    *out = b->groups.size();
    return FULLMAG_FDM_OK;
}

fullmag_fdm_backend *fullmag_fdm_batch_system(fullmag_fdm_batch *b, uint64_t i) {
    return (b && i < b->sys.size()) ? b->sys[i] : nullptr;
}

const char *fullmag_fdm_batch_last_error(const fullmag_fdm_batch *b) {
    return (b && !b->err.empty()) ? b->err.c_str() : nullptr;
}

int fullmag_fdm_batch_advance(fullmag_fdm_batch *b, double t_end, uint64_t max_steps, double dt0,
                              double *out_wall, uint64_t *out_steps) {
    if (!b) return FULLMAG_FDM_ERR_INVALID;
    if (!b->err.empty() || b->sys.empty() || !b->flags) return FULLMAG_FDM_ERR_INVALID;
    if (!(dt0 > 0.0) || !std::isfinite(dt0) || std::isnan(t_end)) {
        b->err = "batch: dt must be positive and finite";
        return FULLMAG_FDM_ERR_INVALID;
    }
    hipSetDevice(b->sys[0]->device);
    const size_t nsys = b->sys.size();
    std::vector<uint64_t> step0(nsys);
    std::vector<double> wall(nsys, -1.0);
    // arm every system's device loop (as fullmag_fdm_backend_advance does)
    for (size_t i = 0; i < nsys; ++i) {
        Backend *h = b->sys[i];
        fm::Ctl &c = *h->ctl_h;
        step0[i] = c.step;
        const double dmin = h->fp64 ? h->ad.dt_min : h->af.dt_min, dmax = h->fp64 ? h->ad.dt_max : h->af.dt_max;
        c.loop = 1;
        c.t_end = t_end;
        c.step_end = sat_add(c.step, max_steps);
        c.dt_carry = dt0;
        double d = std::min(dt0, t_end - c.t);
        if (h->adaptive) d = std::min(std::max(d, dmin), dmax);
        c.dt = d;
        c.done = (c.t >= t_end || max_steps == 0) ? 1 : 0;
        h->obs_fresh = false;
        if (!push_ctl(h) || hipStreamSynchronize(h->stream) != hipSuccess) {
            b->err = "batch: control upload failed: " + h->err;
            return FULLMAG_FDM_ERR_CUDA;
        }
    }
    for (auto &g : b->groups) {
        g.finished = false;
        const bool ok = g.fp64 ? group_prepare<double>(b, g) : group_prepare<float>(b, g);
        if (!ok) return FULLMAG_FDM_ERR_CUDA;
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    int rc = FULLMAG_FDM_OK;
    // every group progresses on its own stream: two chunks in flight, and a
    // group whose older chunk has ended is checked and refilled at once,
    // whatever the other groups are doing (a finished system's launches
    // return at once; at most one chunk after a group's end is wasted)
    auto launch = [&](Batch::Group &g, int slot) {
        return g.fp64 ? group_launch<double>(b, g, slot) : group_launch<float>(b, g, slot);
    };
    for (auto &g : b->groups) {
        g.next = 0;
        if (!launch(g, 0) || !launch(g, 1)) {
            rc = FULLMAG_FDM_ERR_CUDA;
            break;
        }
    }
    while (rc == FULLMAG_FDM_OK) {
        bool active = false, progressed = false;
        for (auto &g : b->groups) {
            if (g.finished) continue;
            active = true;
            const int slot = g.next;
            const hipError_t q = hipEventQuery(g.ev[slot]);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                b->err = "batch: event query failed";
                rc = FULLMAG_FDM_ERR_CUDA;
                break;
            }
            progressed = true;
            const fm::Ctl *f = b->flags + (size_t)slot * nsys;
            bool all = true;
            const double tnow = now();
            for (size_t m : g.members) {
                if (f[m].done) {
                    if (wall[m] < 0) wall[m] = tnow;
                } else {
                    all = false;
                }
            }
            if (all) {
                g.finished = true;
            } else if (!launch(g, slot)) {
                rc = FULLMAG_FDM_ERR_CUDA;
                break;
            }
            g.next = 1 - slot;
        }
        if (!active) break;
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    for (auto &g : b->groups) hipStreamSynchronize(g.stream);
    const hipError_t le = hipGetLastError();
    if (rc == FULLMAG_FDM_OK && le != hipSuccess) {
        b->err = std::string("batch: device error: ") + hipGetErrorString(le);
        rc = FULLMAG_FDM_ERR_CUDA;
    }
    for (size_t i = 0; i < nsys; ++i) {
        Backend *h = b->sys[i];
        if (!pull_ctl(h)) {
            rc = FULLMAG_FDM_ERR_CUDA;
            continue;
        }
        fm::Ctl &c = *h->ctl_h;
        if (out_steps) out_steps[i] = c.step - step0[i];
        if (out_wall) out_wall[i] = wall[i] < 0 ? now() : wall[i];
        c.loop = 0;
        c.done = 0;
        if (!push_ctl(h) || !pull_ctl(h)) rc = FULLMAG_FDM_ERR_CUDA;
    }
    return rc;
}

void fullmag_fdm_batch_destroy(fullmag_fdm_batch *b) {
    if (!b) return;
    for (auto &g : b->groups) {
        if (g.stream) hipStreamSynchronize(g.stream);
        if (g.graph) hipGraphExecDestroy(g.graph);
        if (g.args_d) hipFree(g.args_d);
        for (hipEvent_t e : g.ev)
            if (e) hipEventDestroy(e);
        if (g.stream) hipStreamDestroy(g.stream);
    }
    for (auto *h : b->sys) fullmag_fdm_backend_destroy(h);
    if (b->flags) hipHostFree(b->flags);
    delete b;
}

}  // extern "C"
