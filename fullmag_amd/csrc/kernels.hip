// kernels.hip -- CDNA4 (gfx950) kernels of the FDM micromagnetic step.
//
// One right-hand-side evaluation of the LLG equation (reference chain
// llg_dp45_fp64.cu:185-225: exchange -> demag pack/FFT/multiply/iFFT/unpack
// -> combine -> rhs) is three kernels for a thin film and five for a 3-D
// grid:
//
//   k_fwd_x   Runge-Kutta stage combination y = normalise(m0 + dt sum a k)
//             (reference dp45_rk_stage_*, llg_dp45_fp64.cu:47-133), Ms*y
//             packed two real x-lines per complex sequence, zero-padded R2C
//             Stockham FFT along x in LDS (input half pruned), half spectrum
//             written with y-lines contiguous.
//   k_y       y-lines: pruned forward FFT, symmetric real Newell tensor
//             multiply H = -N M (reference tensor_convolution_fp64_kernel,
//             demag_fp64.cu:151-177), inverse FFT keeping the first ny
//             outputs (thin film); forward / inverse only for 3-D.
//   k_z       3-D only: z-lines forward FFT, 6-component multiply, inverse.
//   k_inv_x   C2R inverse FFT along x for two lines per sequence, 1/P scale
//             (unpack_demag_fft_fp64_kernel, demag_fp64.cu:179-212), then
//             the fused local pass: exchange 6-point stencil with Neumann
//             clamp and inactive-neighbour substitution (exchange_fp64.cu:
//             24-133), Zeeman, uniaxial anisotropy (demag_fp64.cu:214-421),
//             LLG torque (llg_fp64.cu:46-206), and for the last stage the
//             embedded error norm (dp45_error_fp64, llg_dp45_fp64.cu:142-168)
//             and FP64 energy / max-norm partials (reductions_fp64.cu).
//   k_control one work-group: deterministic reduction of the partials, step
//             acceptance, dt control, time/step/parity update on the device
//             (controller semantics of fdm_integrators.rs:647-679).
//
// Without demag the chain is k_stage (combination) + k_local.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include "engine.hpp"

#ifndef FM_REAL
#error "compile once per storage precision with -DFM_REAL=float or -DFM_REAL=double"
#endif

namespace fm {

template <class T>
struct C2 {
    T x, y;
};

// Global-memory accesses.  The buffers reach the kernels through the
// argument record in device memory, so the compiler cannot infer their
// address space and emits FLAT instructions, which also count against the
// LDS wait counter: every wait for an LDS result would then also wait for
// all outstanding memory loads.  These helpers access through address
// space 1 (global) pointers.
template <class T>
using gp = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gp<T> G(T *p) {
    return (gp<T>)p;
}
template <class T>
__device__ __forceinline__ T ldg(const T *p) {
    return *(gp<const T>)p;
}
template <class T>
__device__ __forceinline__ C2<T> ldg(const C2<T> *p) {
    const gp<const T> q = (gp<const T>)p;
    return {q[0], q[1]};
}
template <class T>
__device__ __forceinline__ void stg(T *p, T v) {
    *(gp<T>)p = v;
}
template <class T>
__device__ __forceinline__ void stg(C2<T> *p, C2<T> v) {
    const gp<T> q = (gp<T>)p;
    q[0] = v.x;
    q[1] = v.y;
}

template <class T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by -i (forward) or +i (inverse)
template <class T, bool INV>
__device__ __forceinline__ C2<T> mul_mi(C2<T> a) {
    return INV ? C2<T>{-a.y, a.x} : C2<T>{a.y, -a.x};
}

// Each storage precision is its own translation unit (FM_REAL = float or
// double) with its own copy of the method table.
static __constant__ Method c_methods[M_COUNT];

// Tableau of the step in flight: ABM3 runs its Heun start-up table until
// the history holds three values (Ctl::abm_count changes only in the
// controller, after every kernel of the step has read it).
// Read-only view of the control block through the scalar cache, for the
// kernels that do not write it (every kernel but the controller and the k1
// gate): the scalar cache is invalidated at every launch, and the scalar
// round trip counts on its own counter instead of waiting in line behind the
// vector loads a kernel issues first.
using CtlR = const __attribute__((address_space(4))) Ctl *;
template <class T>
__device__ __forceinline__ CtlR ctl_ro(const Args<T> &A) {
    return (CtlR)A.ctl;
}

// The index is the same for every lane; readfirstlane tells the compiler
// so, and the table fields then come through scalar loads (otherwise the
// control-block read makes the index a vector value, and every field read
// of the table became a dependent vector load with a full memory wait).
template <class T>
__device__ __forceinline__ const Method &method_for(const Args<T> &A) {
    const int idx = A.method == M_ABM3 && ctl_ro(A)->abm_count < 3 ? M_ABM3S : A.method;
    return c_methods[__builtin_amdgcn_readfirstlane(idx)];
}

extern __shared__ __align__(16) unsigned char fm_smem[];

// ------------------------------------------------------------ butterflies --
// Whether a y/z pass reads the control block for its early exit.  The y and
// z passes only transform the spectrum in place: after a device loop has
// finished (or while the k1 gate of a thermal trial is closed) the x passes
// return at once, so nothing reads what a y/z pass leaves, and every trial
// step's first x pass rewrites the spectrum.  Only a batch (finished
// systems would keep their launches busy) and a thermal plan (whose gated
// k1 evaluation is skipped for speed) need the exit; a single system's
// pass then starts its loads without the control block's round trip
// (measured about 2 us of each 512 x 512 film y pass).
template <class T>
__device__ __forceinline__ bool yz_exit_check(const Args<T> &A) {
    return gridDim.y > 1 || A.has_th;
}

template <class T, int R, bool INV>
struct Bfly;

template <class T, bool INV>
struct Bfly<T, 2, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) {
        const C2<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <class T, bool INV>
struct Bfly<T, 4, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) {
        const C2<T> s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
        const C2<T> s1 = cadd(v[1], v[3]), d1 = mul_mi<T, INV>(csub(v[1], v[3]));
        v[0] = cadd(s0, s1);
        v[2] = csub(s0, s1);
        v[1] = cadd(d0, d1);
        v[3] = csub(d0, d1);
    }
};

template <class T, bool INV>
struct Bfly<T, 8, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) {
        // radix-2 split into two radix-4 DFTs of even/odd samples
        C2<T> e[4] = {v[0], v[2], v[4], v[6]};
        C2<T> o[4] = {v[1], v[3], v[5], v[7]};
        Bfly<T, 4, INV>::run(e);
        Bfly<T, 4, INV>::run(o);
        const T h = T(0.70710678118654752440084436210485);
        // o[k] *= W8^k, W8 = exp(-+ 2 pi i / 8)
        o[1] = INV ? C2<T>{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
                   : C2<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
        o[2] = mul_mi<T, INV>(o[2]);
        o[3] = INV ? C2<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
                   : C2<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

template <class T, bool INV>
struct Bfly<T, 3, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) {
        const T c = T(-0.5);
        const T s = INV ? T(0.86602540378443864676372317075294) : T(-0.86602540378443864676372317075294);
        const C2<T> t1 = cadd(v[1], v[2]);
        const C2<T> t2 = csub(v[1], v[2]);
        const C2<T> m = {v[0].x + c * t1.x, v[0].y + c * t1.y};
        // s * i * t2
        const C2<T> r = {-s * t2.y, s * t2.x};
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};

template <class T, bool INV>
struct Bfly<T, 5, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) {
        const T c1 = T(0.30901699437494742410229341718282);
        const T c2 = T(-0.80901699437494742410229341718282);
        const T s1 = T(0.95105651629515357211643933337938);
        const T s2 = T(0.58778525229247312916870595463907);
        const C2<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        const C2<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        const C2<T> p1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
        const C2<T> p2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
        const C2<T> q1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
        const C2<T> q2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
        // forward: X1 = p1 - i q1, X4 = p1 + i q1, X2 = p2 - i q2, X3 = p2 + i q2
        const C2<T> iq1 = mul_mi<T, INV>(q1), iq2 = mul_mi<T, INV>(q2);
        v[0] = {v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y};
        v[1] = cadd(p1, iq1);
        v[4] = csub(p1, iq1);
        v[2] = cadd(p2, iq2);
        v[3] = csub(p2, iq2);
    }
};

}  // namespace fm
#include "regfft.hpp"
namespace fm {

// One Stockham pass (autosort, out of place) with a compile-time radix.
// Sequence sq occupies [sq*n, (sq+1)*n) of src and dst; tw is the LDS copy
// of the forward twiddle table exp(-2 pi i k/n), k < n.  With POW2 every
// index split is a shift/mask (lg_* are the base-2 logarithms).
template <class T, int R, bool INV, bool POW2>
__device__ __forceinline__ void fft_pass(const C2<T> *__restrict__ src, C2<T> *__restrict__ dst, int n, int lg_n,
                                         int ns, int lg_ns, int nseq, const C2<T> *__restrict__ tw, bool first,
                                         int n_in, bool last, int n_out) {
    constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : R == 16 ? 4 : 0;
    const int nr = POW2 ? n >> LR : n / R;
    const int lg_nr = lg_n - LR;
    const int tstep = POW2 ? n >> (lg_ns + LR) : n / (ns * R);
    const int total = nr * nseq;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        int sq, j, k, base;
        if (POW2) {
            sq = b >> lg_nr;
            j = b & (nr - 1);
            k = j & (ns - 1);
            base = ((j >> lg_ns) << (lg_ns + LR)) + k;
        } else {
            sq = b / nr;
            j = b - sq * nr;
            const int q = j / ns;
            k = j - q * ns;
            base = q * ns * R + k;
        }
        const C2<T> *s = src + sq * n;
        C2<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int idx = j + r * nr;
            v[r] = (first && idx >= n_in) ? C2<T>{T(0), T(0)} : s[idx];
        }
        if (ns > 1) {
            const int kt = k * tstep;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C2<T> w = tw[kt * r];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        Bfly<T, R, INV>::run(v);
        C2<T> *d = dst + sq * n;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int o = base + r * ns;
            if (!last || o < n_out) d[o] = v[r];
        }
    }
}

// Generic odd prime radix (R <= kMaxRadix), O(R^2) DFT per butterfly.  The
// inputs are re-read from LDS for every output (no per-thread array, hence no
// scratch memory in any kernel that links the FFT).
template <class T, bool INV>
__device__ void fft_pass_generic(const C2<T> *__restrict__ src, C2<T> *__restrict__ dst, int n, int ns, int nseq,
                                 const C2<T> *__restrict__ tw, bool first, int n_in, bool last, int n_out, int R) {
    const int nr = n / R;
    const int tstep = n / (ns * R);
    const int total = nr * nseq;
    const int wstep = n / R;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        const int sq = b / nr;
        const int j = b - sq * nr;
        const int k = j % ns;
        const C2<T> *sp = src + sq * n;
        const int base = (j / ns) * ns * R + k;
        C2<T> *d = dst + sq * n;
        for (int o = 0; o < R; ++o) {
            C2<T> acc = {T(0), T(0)};
            for (int q = 0; q < R; ++q) {
                const int idx = j + q * nr;
                C2<T> v = (first && idx >= n_in) ? C2<T>{T(0), T(0)} : sp[idx];
                if (ns > 1 && q > 0) {
                    C2<T> w = tw[k * q * tstep];
                    if (INV) w.y = -w.y;
                    v = cmul(v, w);
                }
                C2<T> w = tw[((q * o) % R) * wstep];
                if (INV) w.y = -w.y;
                acc = cadd(acc, cmul(v, w));
            }
            const int oi = base + o * ns;
            if (!last || oi < n_out) d[oi] = acc;
        }
    }
}

// Copies the forward twiddle table of length n into LDS (no barrier: the
// caller's next __syncthreads publishes it).
template <class T>
__device__ __forceinline__ void load_twiddles(C2<T> *lds, const void *gtw, int n) {
    const C2<T> *g = reinterpret_cast<const C2<T> *>(gtw);
    for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g[i];
}

// Batched FFT of `nseq` sequences of length fd.n held in LDS buffer `a`
// (ping-pong with `b`).  Only the first n_in inputs are read (the rest are
// taken as zero) and only the first n_out outputs are written.  Returns the
// buffer that holds the result.  All threads of the block must call it and
// `tw` (LDS) must have been published by a barrier.
template <class T, bool INV>
__device__ C2<T> *lds_fft(C2<T> *a, C2<T> *b, const FftDesc &fd, const C2<T> *__restrict__ tw, int nseq, int n_in,
                          int n_out) {
    const int n = fd.n;
    const bool pow2 = fd.lg >= 0;
    int ns = 1, lg_ns = 0;
    for (int p = 0; p < fd.npass; ++p) {
        const int R = fd.radix[p];
        const bool first = p == 0, last = p == fd.npass - 1;
        if (pow2) {
            switch (R) {
                case 2: fft_pass<T, 2, INV, true>(a, b, n, fd.lg, ns, lg_ns, nseq, tw, first, n_in, last, n_out); break;
                case 4: fft_pass<T, 4, INV, true>(a, b, n, fd.lg, ns, lg_ns, nseq, tw, first, n_in, last, n_out); break;
                default: fft_pass<T, 8, INV, true>(a, b, n, fd.lg, ns, lg_ns, nseq, tw, first, n_in, last, n_out); break;
            }
        } else {
            switch (R) {
                case 2: fft_pass<T, 2, INV, false>(a, b, n, 0, ns, 0, nseq, tw, first, n_in, last, n_out); break;
                case 3: fft_pass<T, 3, INV, false>(a, b, n, 0, ns, 0, nseq, tw, first, n_in, last, n_out); break;
                case 4: fft_pass<T, 4, INV, false>(a, b, n, 0, ns, 0, nseq, tw, first, n_in, last, n_out); break;
                case 5: fft_pass<T, 5, INV, false>(a, b, n, 0, ns, 0, nseq, tw, first, n_in, last, n_out); break;
                case 8: fft_pass<T, 8, INV, false>(a, b, n, 0, ns, 0, nseq, tw, first, n_in, last, n_out); break;
                default: fft_pass_generic<T, INV>(a, b, n, ns, nseq, tw, first, n_in, last, n_out, R); break;
            }
        }
        __syncthreads();
        C2<T> *t = a;
        a = b;
        b = t;
        ns *= R;
        lg_ns += R == 2 ? 1 : R == 4 ? 2 : 3;
    }
    return a;
}

// ------------------------------------------------------------ RK helpers --
template <class T>
__device__ __forceinline__ int kslot(const Args<T> &A, const Method &M, int j, int p) {
    if (M.abm) {
        // ABM3 history ring: f_{n-r} for kidx kHist + r; the predictor's
        // RHS f* goes to the slot of f_{n-2}, which becomes f_n at the push
        if (j >= kHist) return kHistSlot + (ctl_ro(A)->abm_head + 3 - (j - kHist)) % 3;
        if (M.abm == 1 && j == 2) return kHistSlot + (ctl_ro(A)->abm_head + 1) % 3;
    }
    return j == 1 ? p : (j == M.nstage ? 1 - p : j);
}

// ABM3 start-up: the RHS of the new state is also pushed into the history
// (the slot that becomes f_n); NULL otherwise.
template <class T>
__device__ __forceinline__ gp<T> stage_push(const Args<T> &A, const Method &M, int s) {
    if (M.abm != 2 || s != M.nstage) return nullptr;
    return G(A.k_all) + (size_t)(kHistSlot + (ctl_ro(A)->abm_head + 1) % 3) * 3 * (size_t)A.ncell;
}

// Intermediate stage states alternate between two buffers, so that a fused
// kernel can write y_{s+1} while other work-groups still read y_s.
// Offset of the record slot of the step in flight (see Args::rec_step).
template <class T>
__device__ __forceinline__ size_t rec_slot(const Args<T> &A, long long stride) {
    return A.rec_step ? (size_t)((*A.rec_step) % A.rec_seg) * (size_t)stride : 0;
}

// Record offset of stage s's state y_s of the step in flight (s >= 2).  The
// final stage's state is the next step's y_1: it goes to the next slot of
// the ring of rec_seg step slots (wrapping at the ring's end), where that
// step's record starts.
template <class T>
__device__ __forceinline__ size_t rec_y_off(const Args<T> &A, const Method &M, int s) {
    const long long n = *A.rec_step;
    if (s == M.nstage) return (size_t)((n + 1) % A.rec_seg) * (size_t)A.rec_sv;
    return (size_t)(n % A.rec_seg) * (size_t)A.rec_sv + (size_t)(s - 1) * 3 * (size_t)A.ncell;
}

// Record of the field at stage s of the step in flight (NULL when not
// recording); the final stage's field is the next step's stage-1 field.
template <class T>
__device__ __forceinline__ gp<T> rec_h_ptr(const Args<T> &A, const Method &M, int s) {
    if (!A.rec_h) return nullptr;
    const long long n = *A.rec_step;
    if (s == M.nstage) return G(A.rec_h) + (size_t)((n + 1) % A.rec_seg) * (size_t)A.rec_sv;
    return G(A.rec_h) + (size_t)(n % A.rec_seg) * (size_t)A.rec_sv + (size_t)(s - 1) * 3 * (size_t)A.ncell;
}

template <class T>
__device__ __forceinline__ gp<T> stage_buf(const Args<T> &A, int s) {
    return G(A.ybuf) + (size_t)(s & 1) * 3 * (size_t)A.ncell;
}

template <class T>
__device__ __forceinline__ gp<const T> stage_input(const Args<T> &A, const Method &M, int s, int p) {
    const size_t n3 = 3 * (size_t)A.ncell;
    if (s <= 1) return G(A.m_all) + p * n3;
    if (s == M.nstage) return G(A.m_all) + (1 - p) * n3;
    return stage_buf(A, s);
}

// Where stage s writes its state (s >= 2).
template <class T>
__device__ __forceinline__ gp<T> stage_output(const Args<T> &A, const Method &M, int s, int p) {
    return s == M.nstage ? G(A.m_all) + (size_t)(1 - p) * 3 * (size_t)A.ncell : stage_buf(A, s);
}

template <class T>
__device__ __forceinline__ gp<T> stage_rhs(const Args<T> &A, const Method &M, int s, int p) {
    const int slot = s <= 1 ? p : kslot(A, M, s, p);
    return G(A.k_all) + (size_t)slot * 3 * (size_t)A.ncell;
}

// y_s at `cell` (s >= 2 combines and stores; s <= 1 reads m0).
template <class T>
__device__ __forceinline__ void stage_state(const Args<T> &A, const Method &M, int s, int p, double dt, long long cell,
                                            T v[3]) {
    const long long N = A.ncell;
    const gp<const T> m0 = G(A.m_all) + p * 3 * N;
    if (s <= 1) {
        v[0] = m0[cell];
        v[1] = m0[N + cell];
        v[2] = m0[2 * N + cell];
        return;
    }
    // Accumulated in FP64 for both storage precisions: the Dormand-Prince
    // rows have coefficients up to ~12 that cancel to O(1), which costs
    // FP32 several bits; the kernel is memory-bound, so the FP64 ALU work is
    // free.
    const StageRow &R = M.row[s];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, e0 = 0.0, e1 = 0.0, e2 = 0.0;
    const bool efin = A.errv && s == M.nstage;
    for (int q = 0; q < R.nk; ++q) {
        const gp<const T> k = G(A.k_all) + kslot(A, M, R.kidx[q], p) * 3 * N;
        const double w = R.a[q];
        a0 += w * (double)k[cell];
        a1 += w * (double)k[N + cell];
        a2 += w * (double)k[2 * N + cell];
        if (efin) {
            const double we = M.e[q];
            e0 += we * (double)k[cell];
            e1 += we * (double)k[N + cell];
            e2 += we * (double)k[2 * N + cell];
        }
    }
    if (efin) {
        G(A.errv)[cell] = e0;
        G(A.errv)[N + cell] = e1;
        G(A.errv)[2 * N + cell] = e2;
    }
    const double f = dt * R.scale;
    double x = (double)m0[cell] + a0 * f;
    double y = (double)m0[N + cell] + a1 * f;
    double z = (double)m0[2 * N + cell] + a2 * f;
    const double n2 = x * x + y * y + z * z;
    if (n2 > 0.0) {
        const double inv = 1.0 / sqrt(n2);
        x *= inv;
        y *= inv;
        z *= inv;
    } else {
        x = y = z = 0.0;
    }
    const gp<T> out = stage_output(A, M, s, p);
    out[cell] = (T)x;
    out[N + cell] = (T)y;
    out[2 * N + cell] = (T)z;
    if (A.rec_y) {
        T *ry = A.rec_y + rec_y_off(A, M, s);
        ry[cell] = (T)x;
        ry[N + cell] = (T)y;
        ry[2 * N + cell] = (T)z;
        G(A.rec_vn)[rec_slot(A, A.rec_ss) + (size_t)(s - 2) * N + cell] = (T)sqrt(n2);
    }
    v[0] = (T)x;
    v[1] = (T)y;
    v[2] = (T)z;
}

struct Acc {
    double v[kStatSlots];
    __device__ void init() {
#pragma unroll
        for (int i = 0; i < kStatSlots; ++i) v[i] = 0.0;
    }
};

__device__ __forceinline__ bool slot_is_max(int i) {
    return i == ST_ERR || i == ST_HEFF2 || i == ST_HD2 || i == ST_RHS2;
}

// Block-wide reduction of Acc in a fixed order (deterministic): thread
// i < kStatSlots returns the total of slot i, the others 0.  Must be called
// by all threads.  The two stores below keep the destination's address
// space explicit: per-work-group partials go to global memory, the
// controller's totals to LDS.
__device__ double block_reduce_slot(Acc &a) {
    __shared__ double red[kStatSlots][16];  // up to 1024 threads
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // wave butterfly, all slots per exchange distance: the slots' lane
    // exchanges are independent, so each distance costs one exchange
    // latency instead of one per slot (the per-slot order of operations,
    // and so the result, is that of a slot-at-a-time butterfly)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double y[kStatSlots];
#pragma unroll
        for (int i = 0; i < kStatSlots; ++i) y[i] = __shfl_xor(a.v[i], off, 64);
#pragma unroll
        for (int i = 0; i < kStatSlots; ++i) a.v[i] = slot_is_max(i) ? fmax(a.v[i], y[i]) : a.v[i] + y[i];
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < kStatSlots; ++i) red[i][wid] = a.v[i];
    }
    __syncthreads();
    double x = 0.0;
    if (threadIdx.x < kStatSlots) {
        const int i = threadIdx.x;
        x = red[i][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) x = slot_is_max(i) ? fmax(x, red[i][w]) : x + red[i][w];
    }
    return x;
}

// Totals into a work-group's partial record (device memory).
__device__ void block_reduce_global(Acc &a, double *out) {
    const double x = block_reduce_slot(a);
    if (threadIdx.x < kStatSlots) G(out)[threadIdx.x] = x;
}

// Totals into an LDS array of the calling kernel.
__device__ void block_reduce_lds(Acc &a, double (&tot)[kStatSlots]) {
    const double x = block_reduce_slot(a);
    if (threadIdx.x < kStatSlots) tot[threadIdx.x] = x;
}

// Waveform value f(t) of the time-dependent Zeeman term.
template <class T>
__device__ __forceinline__ T wave_factor(const Args<T> &A, double t) {
    switch (A.wave_kind) {
        case WAVE_RAMP: {
            const double u = (t - A.wave_t0) / (A.wave_t1 - A.wave_t0);
            return T(u < 0.0 ? 0.0 : u > 1.0 ? 1.0 : u);
        }
        case WAVE_SINC: {
            const double x = 2.0 * 3.14159265358979323846 * A.wave_freq * (t - A.wave_t0);
            return T(x == 0.0 ? 1.0 : sin(x) / x);
        }
        case WAVE_SINE: return T(sin(2.0 * 3.14159265358979323846 * A.wave_freq * t + A.wave_phase));
        default: return T(0);
    }
}

// Thermal field of one cell, fdm_fields.rs:818-905: the SplitMix64
// finaliser of a key of (seed, counter, cell, stream) gives four uniforms in
// (0, 1], Box-Muller three Gaussians.  The counter is the number of accepted
// steps (one per step call, shared by all stages and attempts).
__device__ __forceinline__ unsigned long long th_mix(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double th_uniform(unsigned long long seed, unsigned long long step, unsigned long long cell,
                                             unsigned long long stream) {
    const unsigned long long key = seed + step * 0x9E3779B97F4A7C15ull + cell * 0x517CC1B727220A95ull +
                                   stream * 0x6C62272E07BB0142ull;
    return ((double)(th_mix(key) >> 11) + 1.0) / ((double)(1ull << 53) + 1.0);
}
template <class T>
__device__ __forceinline__ void thermal_field(const Args<T> &A, int cell, T ms_i, T h[3]) {
    const double dt = A.th_dt > 0.0 ? A.th_dt : ctl_ro(A)->dt;
    if (!(dt > 0.0)) return;
    const double sig = sqrt(A.th_pre / ((double)ms_i * dt));
    const unsigned long long st = ctl_ro(A)->step, c = (unsigned long long)cell;
    const double u1 = fmax(th_uniform(A.th_seed, st, c, 0), 1e-300), u2 = th_uniform(A.th_seed, st, c, 1);
    const double u3 = fmax(th_uniform(A.th_seed, st, c, 2), 1e-300), u4 = th_uniform(A.th_seed, st, c, 3);
    const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    const double tp = 2.0 * 3.14159265358979323846;
    h[0] += (T)(sig * r1 * cos(tp * u2));
    h[1] += (T)(sig * r1 * sin(tp * u2));
    h[2] += (T)(sig * r2 * cos(tp * u4));
}

// Time-dependent factors of one stage: Zeeman waveform and Oersted current
// envelope (native/backends/fdm/src/demag_fp64.cu:628-650 forms).
template <class T>
struct TimeF {
    T wave, oe;
};
template <class T>
__device__ __forceinline__ TimeF<T> time_factors(const Args<T> &A, double t) {
    TimeF<T> f;
    f.wave = A.wave_kind ? wave_factor(A, t) : T(0);
    double e = 1.0;
    if (A.oe_kind == 1) e = sin(2.0 * 3.14159265358979323846 * A.oe_freq * t + A.oe_phase) + A.oe_offset;
    else if (A.oe_kind == 2) e = (t >= A.oe_t_on && t < A.oe_t_off) ? 1.0 : 0.0;
    f.oe = T(e);
    return f;
}

// Evaluation time of stage s of the trial step from the control block.
template <class T, class CP>
__device__ __forceinline__ double stage_time(const Args<T> &A, const Method &M, int s, CP ctl, double dt) {
    return s <= 1 ? ctl->t : ctl->t + M.node[s] * dt;
}

// Everything the local pass reads for one cell.  The pass is bound by memory
// latency, so the loads of a thread's cells are issued together (gather) and
// only then consumed (compute); no load depends on another load's value
// except the optional region-table lookup.
template <class T>
struct CellIn {
    T m[3];           // state at the cell
    T v[6][3];        // box-clamped neighbour states (-x,+x,-y,+y,-z,+z)
    T ms, ku1, ku2, ax[3];
    T kc[3];          // cubic constants (further terms)
    T oe[3];          // Oersted field at the full current (further terms)
    T bw[6];          // boundary-tier exchange weights (further terms)
    T phi;            // volume fraction (further terms)
    T hc[3];          // sparse demag correction (further terms)
    T m0[3];          // state at the start of the step (mixed tolerance)
    double ev[3];     // sum of the error terms before the final stage (Args::errv)
    unsigned char act, nact[6], wm;
    uint32_t rc, rn[6];
};

template <class T, int FT>
__device__ __forceinline__ void gather_cell(const Args<T> &A, const Method &M, int p, int cell, int x, int y, int z,
                                            gp<const T> __restrict__ yin, bool err, CellIn<T> &g) {
    constexpr bool LUT = (FT & 1) != 0, XT = (FT & 2) != 0;
    const int N = (int)A.ncell;
    const int sy = A.nx, sz = A.nx * A.ny;
    int nb[6] = {x > 0 ? cell - 1 : cell, x < A.nx - 1 ? cell + 1 : cell,
                 y > 0 ? cell - sy : cell, y < A.ny - 1 ? cell + sy : cell,
                 z > 0 ? cell - sz : cell, z < A.nz - 1 ? cell + sz : cell};
    if constexpr (XT) {
        // periodic axes wrap (neighbor_index, fdm_types.rs:119-132)
        if (A.pbc & 1) {
            if (x == 0) nb[0] = cell + A.nx - 1;
            if (x == A.nx - 1) nb[1] = cell - (A.nx - 1);
        }
        if (A.pbc & 2) {
            if (y == 0) nb[2] = cell + (A.ny - 1) * sy;
            if (y == A.ny - 1) nb[3] = cell - (A.ny - 1) * sy;
        }
        if (A.pbc & 4) {
            if (z == 0) nb[4] = cell + (A.nz - 1) * sz;
            if (z == A.nz - 1) nb[5] = cell - (A.nz - 1) * sz;
        }
        // multilayer: the planes are separate bodies (no z neighbours)
        if (A.layered) nb[4] = nb[5] = cell;
    }
    // activity of the cell and of its six stencil neighbours in one byte
    // (one load instead of seven dependent ones)
    const unsigned bits = A.nmask ? (unsigned)G(A.nmask)[cell] : 0x7Fu;
    g.act = bits & 1u;
#pragma unroll
    for (int q = 0; q < 6; ++q) g.nact[q] = (bits >> (q + 1)) & 1u;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        g.m[c] = yin[c * N + cell];
#pragma unroll
        for (int q = 0; q < 6; ++q) g.v[q][c] = yin[c * N + nb[q]];
    }
    g.ms = A.ms_cell ? G(A.ms_cell)[cell] : A.ms;
    if (A.has_uni) {
        g.ku1 = A.ku1_cell ? G(A.ku1_cell)[cell] : A.ku1_u;
        g.ku2 = A.ku2_cell ? G(A.ku2_cell)[cell] : A.ku2_u;
#pragma unroll
        for (int c = 0; c < 3; ++c) g.ax[c] = A.axis_cell ? G(A.axis_cell)[c * N + cell] : A.axis[c];
    }
    g.wm = A.wave_mask ? G(A.wave_mask)[cell] : 1;
    if constexpr (XT) {
        if (A.has_cub) {
            g.kc[0] = A.kc1_cell ? G(A.kc1_cell)[cell] : A.kc1_u;
            g.kc[1] = A.kc2_cell ? G(A.kc2_cell)[cell] : A.kc2_u;
            g.kc[2] = A.kc3_cell ? G(A.kc3_cell)[cell] : A.kc3_u;
        }
        if (A.oe_field) {
#pragma unroll
            for (int c = 0; c < 3; ++c) g.oe[c] = G(A.oe_field)[c * N + cell];
        }
        if (A.bc_w) {
#pragma unroll
            for (int q = 0; q < 6; ++q) g.bw[q] = G(A.bc_w)[q * N + cell];
        }
        g.phi = A.bc_phi ? G(A.bc_phi)[cell] : T(1);
        g.hc[0] = g.hc[1] = g.hc[2] = T(0);
        if (A.dc_ptr) {
            // demag_boundary_fp64.cu:23-76: H -= dN . (phi Ms y) over the
            // cell's correction sources
            const int b = G(A.dc_ptr)[cell], e = G(A.dc_ptr)[cell + 1];
            for (int j = b; j < e; ++j) {
                const int sc = G(A.dc_src)[j];
                const T w = G(A.ms_pack)[sc];
                const T M0 = w * yin[sc], M1 = w * yin[N + sc], M2 = w * yin[2 * N + sc];
                const gp<const T> t = G(A.dc_n) + 6 * (size_t)j;
                g.hc[0] -= t[0] * M0 + t[1] * M1 + t[2] * M2;
                g.hc[1] -= t[1] * M0 + t[3] * M1 + t[4] * M2;
                g.hc[2] -= t[2] * M0 + t[4] * M1 + t[5] * M2;
            }
        }
    }
    if (LUT) {
        g.rc = G(A.region)[cell];
#pragma unroll
        for (int q = 0; q < 6; ++q) g.rn[q] = G(A.region)[nb[q]];
    }
    if (err) {
        // every adaptive tableau has Method::efuse: the final stage
        // combination left the other error terms' sum in errv
#pragma unroll
        for (int c = 0; c < 3; ++c) g.ev[c] = G(A.errv)[c * N + cell];
        if (A.rtol > 0.0) {
            const gp<const T> m0 = G(A.m_all) + (size_t)p * 3 * N;
#pragma unroll
            for (int c = 0; c < 3; ++c) g.m0[c] = m0[c * N + cell];
        }
    }
}

// Local field terms, LLG torque, and (optionally) statistics for one cell.
template <class T, int FT>
__device__ __forceinline__ void compute_cell(const Args<T> &A, const Method &M, int cell, double dt,
                                             const CellIn<T> &g, gp<T> __restrict__ kout, const T hd_in[3], TimeF<T> tf,
                                             bool stats, bool err, Acc &acc, gp<T> __restrict__ hrec,
                                             T kret[3] = nullptr, gp<T> __restrict__ kpush = nullptr) {
    constexpr bool LUT = (FT & 1) != 0, XT = (FT & 2) != 0;
    const T wf = tf.wave;
    const int N = (int)A.ncell;
    if (!g.act) {
        kout[cell] = T(0);
        kout[N + cell] = T(0);
        kout[2 * N + cell] = T(0);
        if (kpush) kpush[cell] = kpush[N + cell] = kpush[2 * N + cell] = T(0);
        if (kret) kret[0] = kret[1] = kret[2] = T(0);
        if (hrec) hrec[cell] = hrec[N + cell] = hrec[2 * N + cell] = T(0);
        if (A.obs_hex) {
            G(A.obs_hex)[cell] = G(A.obs_hex)[N + cell] = G(A.obs_hex)[2 * N + cell] = T(0);
            G(A.obs_hd)[cell] = G(A.obs_hd)[N + cell] = G(A.obs_hd)[2 * N + cell] = T(0);
            G(A.obs_heff)[cell] = G(A.obs_heff)[N + cell] = G(A.obs_heff)[2 * N + cell] = T(0);
        }
        return;
    }
    const T *m = g.m;
    const T ms_i = g.ms;
    T hex[3];
    bool tiered = false;
    if constexpr (XT) tiered = A.bc_w != nullptr;
    if (tiered) {
        // boundary-corrected stencils (exchange_t0_fp64.cu:52-165,
        // exchange_t1_fp64.cu:62-198): per-face weights prepared on the host
        // from the face links or boundary distances; neighbours as stored
        const T ww[6] = {A.idx2, A.idx2, A.idy2, A.idy2, A.idz2, A.idz2};
        T aw[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) aw[q] = g.bw[q] * ww[q] * (LUT ? G(A.lut)[g.rc * 256u + g.rn[q]] : T(1));
        const T pre = LUT ? (A.ms_cell ? A.two_mu0inv / ms_i : A.lut_pre) : (A.ms_cell ? A.two_a_mu0 / ms_i : A.ex_pre);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T acc2 = T(0);
#pragma unroll
            for (int q = 0; q < 6; ++q) acc2 += aw[q] * (g.v[q][c] - m[c]);
            hex[c] = pre * acc2;
        }
    } else if (!LUT) {
        const T pre = A.ms_cell ? A.two_a_mu0 / ms_i : A.ex_pre;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T w[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) w[q] = g.nact[q] ? g.v[q][c] : m[c];
            const T mc = m[c];
            // neighbour differences first: each is exact for neighbours
            // within a factor of two of the cell's value (Sterbenz), so the
            // second difference of a smooth field keeps its digits in FP32
            // (w1 - 2m + w0 rounds 2m - w1 to ~1e-7 absolute, a large part
            // of a long spin wave's second difference)
            const T lap = ((w[1] - mc) + (w[0] - mc)) * A.idx2 + ((w[3] - mc) + (w[2] - mc)) * A.idy2 +
                          ((w[5] - mc) + (w[4] - mc)) * A.idz2;
            hex[c] = pre * lap;
        }
    } else {
        T aij[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) aij[q] = G(A.lut)[g.rc * 256u + g.rn[q]];
        const T ww[6] = {A.idx2, A.idx2, A.idy2, A.idy2, A.idz2, A.idz2};
        const T pre = A.ms_cell ? A.two_mu0inv / ms_i : A.lut_pre;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T acc2 = T(0);
#pragma unroll
            for (int q = 0; q < 6; ++q) acc2 += aij[q] * ((g.nact[q] ? g.v[q][c] : m[c]) - m[c]) * ww[q];
            hex[c] = pre * acc2;
        }
    }
    T hext[3] = {T(0), T(0), T(0)};
    if (A.has_ext) {
        hext[0] = A.hext[0];
        hext[1] = A.hext[1];
        hext[2] = A.hext[2];
    }
    if (A.wave_kind && g.wm) {
        hext[0] += wf * T(A.wave_amp[0]);
        hext[1] += wf * T(A.wave_amp[1]);
        hext[2] += wf * T(A.wave_amp[2]);
    }
    T hani[3] = {T(0), T(0), T(0)};
    if (A.has_uni) {
        T c1 = A.uni_c1, c2 = A.uni_c2;
        if (A.ku1_cell || A.ku2_cell || A.ms_cell) {
            c1 = A.two_mu0inv * g.ku1 / ms_i;
            c2 = T(2) * A.two_mu0inv * g.ku2 / ms_i;
        }
        const T mu = m[0] * g.ax[0] + m[1] * g.ax[1] + m[2] * g.ax[2];
        const T coef = c1 * mu + c2 * mu * mu * mu;
        hani[0] = g.ax[0] * coef;
        hani[1] = g.ax[1] * coef;
        hani[2] = g.ax[2] * coef;
    }
    T hd[3] = {hd_in[0], hd_in[1], hd_in[2]};
    if constexpr (XT) {
#pragma unroll
        for (int c = 0; c < 3; ++c) hd[c] += g.hc[c];
    }
    T h[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) h[c] = hex[c] + hd[c] + hext[c] + hani[c];
    // further terms: cubic anisotropy (fdm_fields.rs:648-705, Kc3 as
    // demag_fp64.cu:301-330), magnetoelastic (magnetoelastic.rs:52-75), DMI
    // over the box-clamped neighbour values as stored (fdm_fields.rs:707-816),
    // Oersted field times the current envelope
    T hcub[3] = {T(0), T(0), T(0)}, hdmi[3] = {T(0), T(0), T(0)};
    double e_mel = 0.0;
    if constexpr (XT) {
        const T inv_ms = T(1) / ms_i;
        if (A.has_cub) {
            const T *c1 = A.cub_ax, *c2 = A.cub_ax + 3, *c3 = A.cub_ax + 6;
            const T m1 = m[0] * c1[0] + m[1] * c1[1] + m[2] * c1[2];
            const T m2 = m[0] * c2[0] + m[1] * c2[1] + m[2] * c2[2];
            const T m3 = m[0] * c3[0] + m[1] * c3[1] + m[2] * c3[2];
            const T a1 = m1 * m1, a2 = m2 * m2, a3 = m3 * m3;
            const T sig = a1 * a2 + a2 * a3 + a1 * a3;
            const T pf = A.two_mu0inv * inv_ms, pf3 = T(2) * pf * g.kc[2] * sig;
            const T g1 = -(pf * (g.kc[0] * m1 * (a2 + a3) + g.kc[1] * m1 * a2 * a3) + pf3 * m1 * (a2 + a3));
            const T g2 = -(pf * (g.kc[0] * m2 * (a1 + a3) + g.kc[1] * m2 * a1 * a3) + pf3 * m2 * (a1 + a3));
            const T g3 = -(pf * (g.kc[0] * m3 * (a1 + a2) + g.kc[1] * m3 * a1 * a2) + pf3 * m3 * (a1 + a2));
#pragma unroll
            for (int c = 0; c < 3; ++c) hcub[c] = g1 * c1[c] + g2 * c2[c] + g3 * c3[c];
        }
        if (A.has_mel) {
            const T *e = A.mel_e;  // e11 e22 e33 e23 e13 e12
            const T pf = -T(0.5) * A.two_mu0inv * inv_ms, b1 = T(2) * A.mel_b1, b2 = T(2) * A.mel_b2;
            h[0] += pf * (b1 * m[0] * e[0] + b2 * (m[1] * e[5] + m[2] * e[4]));
            h[1] += pf * (b1 * m[1] * e[1] + b2 * (m[0] * e[5] + m[2] * e[3]));
            h[2] += pf * (b1 * m[2] * e[2] + b2 * (m[0] * e[4] + m[1] * e[3]));
            if (stats) {
                const double mx = m[0], my = m[1], mz = m[2];
                e_mel = A.cell_vol * ((double)A.mel_b1 * (mx * mx * e[0] + my * my * e[1] + mz * mz * e[2]) +
                                      2.0 * (double)A.mel_b2 * (mx * my * e[5] + mx * mz * e[4] + my * mz * e[3]));
            }
        }
        if (A.dmi_i != T(0)) {
            const T pf = A.two_mu0inv * A.dmi_i * inv_ms;
            hdmi[0] += pf * (g.v[1][2] - g.v[0][2]) * A.i2x;
            hdmi[1] += pf * (g.v[3][2] - g.v[2][2]) * A.i2y;
            hdmi[2] -= pf * ((g.v[1][0] - g.v[0][0]) * A.i2x + (g.v[3][1] - g.v[2][1]) * A.i2y);
        }
        if (A.dmi_b != T(0)) {
            const T pf = -A.two_mu0inv * A.dmi_b * inv_ms;
            hdmi[0] += pf * ((g.v[3][2] - g.v[2][2]) * A.i2y - (g.v[5][1] - g.v[4][1]) * A.i2z);
            hdmi[1] += pf * ((g.v[5][0] - g.v[4][0]) * A.i2z - (g.v[1][2] - g.v[0][2]) * A.i2x);
            hdmi[2] += pf * ((g.v[1][1] - g.v[0][1]) * A.i2x - (g.v[3][0] - g.v[2][0]) * A.i2y);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) h[c] += hcub[c] + hdmi[c] + (A.oe_field ? tf.oe * g.oe[c] : T(0));
        if (A.has_th) thermal_field(A, cell, ms_i, h);
    }
    // LLG: k = -gb (m x H + alpha m x (m x H))
    const T pr[3] = {m[1] * h[2] - m[2] * h[1], m[2] * h[0] - m[0] * h[2], m[0] * h[1] - m[1] * h[0]};
    const T dm[3] = {m[1] * pr[2] - m[2] * pr[1], m[2] * pr[0] - m[0] * pr[2], m[0] * pr[1] - m[1] * pr[0]};
    T k[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) k[c] = -A.gamma_bar * ((A.precession ? pr[c] : T(0)) + A.alpha * dm[c]);
    // torques added to every RHS (native/backends/fdm/src/llg_fp64.cu:58-205;
    // forms of fdm_fields.rs:1446-1680)
    if constexpr (XT) {
        const T inv_ms = T(1) / ms_i;
        auto cross = [](const T *a, const T *b, T *o) {
            o[0] = a[1] * b[2] - a[2] * b[1];
            o[1] = a[2] * b[0] - a[0] * b[2];
            o[2] = a[0] * b[1] - a[1] * b[0];
        };
        if (A.has_zl) {
            // upwind (u.grad) m; a box face gives a zero difference, as the
            // reference's skipped term
            T du[3] = {T(0), T(0), T(0)};
            const T inv_d[3] = {A.i1x, A.i1y, A.i1z};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const T u = A.zl_u[a] * inv_ms;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    du[c] += u > T(0) ? u * (m[c] - g.v[2 * a][c]) * inv_d[a]
                                      : (u < T(0) ? u * (g.v[2 * a + 1][c] - m[c]) * inv_d[a] : T(0));
            }
            T c1[3], c2[3];
            cross(m, du, c1);
            cross(m, c1, c2);
#pragma unroll
            for (int c = 0; c < 3; ++c) k[c] += -c2[c] - A.zl_beta * c1[c];
        }
        if (A.has_slon) {
            const T mp_dot = m[0] * A.sl_p[0] + m[1] * A.sl_p[1] + m[2] * A.sl_p[2];
            const T gg = A.sl_deg * A.sl_l2 / ((A.sl_l2 + T(1)) + (A.sl_l2 - T(1)) * mp_dot);
            const T b = A.sl_pf * inv_ms * gg;
            T c1[3], c2[3];
            cross(m, A.sl_p, c1);
            cross(m, c1, c2);
#pragma unroll
            for (int c = 0; c < 3; ++c) k[c] += b * (c2[c] + A.sl_eps * c1[c]);
        }
        if (A.has_sot) {
            const T amp = A.sot_pf * inv_ms;
            T c1[3], c2[3];
            cross(m, A.sot_s, c1);
            cross(m, c1, c2);
#pragma unroll
            for (int c = 0; c < 3; ++c) k[c] += amp * (-A.sot_dl * c2[c] + A.sot_fl * c1[c]);
        }
    }
    kout[cell] = k[0];
    kout[N + cell] = k[1];
    kout[2 * N + cell] = k[2];
    if (kpush) {
        kpush[cell] = k[0];
        kpush[N + cell] = k[1];
        kpush[2 * N + cell] = k[2];
    }
    if (kret) {
        kret[0] = k[0];
        kret[1] = k[1];
        kret[2] = k[2];
    }
    if (hrec) {
        hrec[cell] = h[0];
        hrec[N + cell] = h[1];
        hrec[2 * N + cell] = h[2];
    }
    if (A.obs_hex) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            G(A.obs_hex)[c * N + cell] = hex[c];
            G(A.obs_hd)[c * N + cell] = hd[c];
            G(A.obs_heff)[c * N + cell] = h[c];
        }
    }
    if (stats) {
        const double md[3] = {(double)m[0], (double)m[1], (double)m[2]};
        auto dotd = [&](const T *f) { return md[0] * (double)f[0] + md[1] * (double)f[1] + md[2] * (double)f[2]; };
        auto sq = [](const T *f) {
            return (double)f[0] * (double)f[0] + (double)f[1] * (double)f[1] + (double)f[2] * (double)f[2];
        };
        double ep = A.ms_cell ? A.mu0_vol * (double)ms_i : A.e_pre;
        if constexpr (XT) ep *= (double)g.phi;  // boundary tier: cell volume phi V
        acc.v[ST_EEX] += -0.5 * ep * dotd(hex);
        acc.v[ST_EDEMAG] += -0.5 * ep * dotd(hd);
        acc.v[ST_EEXT] += -1.0 * ep * dotd(hext);
        acc.v[ST_EANI] += -0.5 * ep * dotd(hani);
        if constexpr (XT) {
            acc.v[ST_ECUB] += -0.5 * ep * dotd(hcub);
            acc.v[ST_EDMI] += -0.5 * ep * dotd(hdmi);
            acc.v[ST_EMEL] += e_mel * (double)g.phi;
        }
        acc.v[ST_HEFF2] = fmax(acc.v[ST_HEFF2], sq(h));
        acc.v[ST_HD2] = fmax(acc.v[ST_HD2], sq(hd));
        acc.v[ST_RHS2] = fmax(acc.v[ST_RHS2], sq(k));
    }
    if (err) {
        // terms in tableau order: the earlier ones were summed by the final
        // stage combination, the final stage's k is the last term
        const double w = M.e[M.nerr - 1];
        double e0 = g.ev[0] + w * (double)k[0];
        double e1 = g.ev[1] + w * (double)k[1];
        double e2 = g.ev[2] + w * (double)k[2];
        e0 *= dt;
        e1 *= dt;
        e2 *= dt;
        double v = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
        if (A.rtol > 0.0) {
            const double a = (double)g.m0[0], b = (double)g.m0[1], c = (double)g.m0[2];
            v /= (A.atol + A.rtol * fmax(sqrt(a * a + b * b + c * c), 1e-30));
        }
        acc.v[ST_ERR] = fmax(acc.v[ST_ERR], v);
    }
}

// Cells per thread whose loads are in flight together.
template <class T>
struct Chunk {
    static constexpr int value = sizeof(T) == 4 ? 4 : 2;
};

// Logical work-group index that keeps consecutive logical groups on one XCD:
// the hardware deals work-groups round-robin over the 8 XCDs (each with its
// own L2), so neighbouring x-lines, which share spectrum cache lines, would
// otherwise be fetched by 8 different L2s.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned g) {
    const unsigned x = b & 7u, j = b >> 3;
    const unsigned q = g >> 3, r = g & 7u;
    return x * q + (x < r ? x : r) + j;
}

// Global-memory index of the four-step twiddle W_N^{t k2} (stored at
// [k2*R1 + t]) that the forward step's thread t applies to its value k2.
// For R1 = R2 the table is symmetric and [t*R1 + k2] holds the same value
// with a thread's k2 consecutive, so its loads merge into 16-byte ones
// (x passes 512 x 512 film: first pass 10.7 -> 9.8 us FP32; the film y
// pass, whose lanes then read 32 lines per load, measured slower: 10.8 ->
// 13.8 us, and keeps [k2*R1 + t]).  LDS reads keep [k2*R1 + t]:
// consecutive lanes there hit different banks.
template <int R1, int R2>
__device__ __forceinline__ int tw_fwd_gidx(int t, int k2) {
    if constexpr (R1 == R2) return t * R1 + k2;
    else return k2 * R1 + t;
}

// Register four-step FFT of `nseq` sequences of length N = R1*R2 held in
// natural order in LDS `buf` (sequence stride sb >= N), in place.  Thread t1 < R1 of a
// sequence holds x[t1 + R1 n2] (inputs n >= n_in are zero), does the
// R2-point DFT in VGPRs, applies W_N^{t1 k2} (tw4[k2*R1 + t1]) and writes a
// padded transpose to `tb`; thread k2 < R2 then does the R1-point DFT and
// writes X[k2 + R2 k1] back to `buf`.  `tb` holds nseq*TB elements.  All
// threads of the block must call it (it contains barriers).
template <int R1, int R2>
struct FourStep {
    static constexpr int TS = R1 > R2 ? R1 : R2;
    static constexpr int N = R1 * R2;
    static constexpr int P1 = R1 + 1, P2 = R2 + 1;
    static constexpr int TB = R2 * P1 > R1 * P2 ? R2 * P1 : R1 * P2;
};

// Exchange with the neighbouring lane of a pair (lanes 2j, 2j+1) through a
// DPP quad permutation [1,0,3,2] (a VALU move, no LDS round trip).
__device__ __forceinline__ float pair_swap(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double pair_swap(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// R-point DFT shared by a lane pair: lane e of the pair holds x[e + 2n]
// (n < R/2) and returns X[k + e R/2] (k < R/2): an R/2-point DFT per lane,
// the twiddle W_R^{e k} and one radix-2 butterfly across the pair.  Both
// lanes of a pair must be active.
template <class T, int R, bool INV>
__device__ __forceinline__ void dft_pair(C2<T> *v, int e) {
    constexpr int Q = R / 2;
    DftReg<T, Q, INV>::run(v);
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        const T c = T(TwHolder<R>::tw.c[k]);
        const T sn = INV ? T(TwHolder<R>::tw.s[k]) : T(-TwHolder<R>::tw.s[k]);
        const C2<T> w = cmul(v[k], C2<T>{c, sn});
        v[k] = e ? w : v[k];
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const C2<T> o = {pair_swap(v[k].x), pair_swap(v[k].y)};
        v[k] = e ? csub(o, v[k]) : cadd(v[k], o);
    }
}

// One lane per line DFT (FP64: the pair exchange doubles the register
// traffic of the FP64 values and measured slower).
template <class T, int R1, int R2, bool TL>
__device__ void fs_forward1(C2<T> *buf, C2<T> *tb, const C2<T> *__restrict__ tw4, int nseq, int n_in, int sb) {
    using F = FourStep<R1, R2>;
    const int spb = blockDim.x / F::TS;  // sequences per round
    for (int base = 0; base < nseq; base += spb) {
        const int sq = base + threadIdx.x / F::TS, t = threadIdx.x % F::TS;
        const bool on = sq < nseq && (int)threadIdx.x < spb * F::TS;
        C2<T> *b = buf + sq * sb;
        C2<T> *tt = tb + (sq - base) * F::TB;
        if (on && t < R1) {
            C2<T> v[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) {
                const int n = t + R1 * n2;
                v[n2] = n < n_in ? b[n] : C2<T>{T(0), T(0)};
            }
            DftReg<T, R2, false>::run(v);
#pragma unroll
            for (int k2 = 1; k2 < R2; ++k2)
                v[k2] = cmul(v[k2], TL ? tw4[k2 * R1 + t] : ldg(tw4 + tw_fwd_gidx<R1, R2>(t, k2)));
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) tt[k2 * F::P1 + t] = v[k2];
        }
        __syncthreads();
        if (on && t < R2) {
            C2<T> u[R1];
#pragma unroll
            for (int t1 = 0; t1 < R1; ++t1) u[t1] = tt[t * F::P1 + t1];
            DftReg<T, R1, false>::run(u);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) b[t + R2 * k1] = u[k1];
        }
        __syncthreads();
    }
}

template <class T, int R1, int R2, bool TL>
__device__ void fs_inverse1(C2<T> *buf, C2<T> *tb, const C2<T> *__restrict__ tw4, int nseq, int n_out, int sb) {
    using F = FourStep<R1, R2>;
    const int spb = blockDim.x / F::TS;
    for (int base = 0; base < nseq; base += spb) {
        const int sq = base + threadIdx.x / F::TS, t = threadIdx.x % F::TS;
        const bool on = sq < nseq && (int)threadIdx.x < spb * F::TS;
        C2<T> *b = buf + sq * sb;
        C2<T> *tt = tb + (sq - base) * F::TB;
        if (on && t < R2) {
            C2<T> u[R1];
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) u[k1] = b[t + R2 * k1];
            DftReg<T, R1, true>::run(u);
#pragma unroll
            for (int t1 = 1; t1 < R1; ++t1) {
                C2<T> w = TL ? tw4[t * R1 + t1] : ldg(tw4 + (t * R1 + t1));
                w.y = -w.y;
                u[t1] = cmul(u[t1], w);
            }
#pragma unroll
            for (int t1 = 0; t1 < R1; ++t1) tt[t1 * F::P2 + t] = u[t1];
        }
        __syncthreads();
        if (on && t < R1) {
            C2<T> v[R2];
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[k2] = tt[t * F::P2 + k2];
            DftReg<T, R2, true>::run(v);
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) {
                const int n = t + R1 * n2;
                if (n < n_out) b[n] = v[n2];
            }
        }
        __syncthreads();
    }
}

// Register four-step FFT, two lanes per line DFT in FP32 (64 lanes per
// sequence at R1 = R2 = 32, so three sequences keep three of a
// work-group's four SIMDs busy; the x passes of the 512 x 512 film 20-30 %
// faster); FP64 keeps one lane per line DFT.  TL: tw4 is in LDS (else
// global memory).
template <class T, int R1, int R2, bool TL = false>
__device__ void fs_forward(C2<T> *buf, C2<T> *tb, const C2<T> *__restrict__ tw4, int nseq, int n_in, int sb) {
    if constexpr (sizeof(T) == 8) {
        fs_forward1<T, R1, R2, TL>(buf, tb, tw4, nseq, n_in, sb);
        return;
    }
    using F = FourStep<R1, R2>;
    constexpr int L = 2 * F::TS;               // lanes per sequence
    const int spb = blockDim.x / L;            // sequences per round
    for (int base = 0; base < nseq; base += spb) {
        const int sq = base + threadIdx.x / L, r = threadIdx.x % L, t = r >> 1, e = r & 1;
        const bool on = sq < nseq && (int)threadIdx.x < spb * L;
        C2<T> *b = buf + sq * sb;
        C2<T> *tt = tb + (sq - base) * F::TB;
        if (on && t < R1) {
            // R2-point DFT over n2 of x[t + R1 n2]: this lane takes n2 = e + 2n
            C2<T> v[R2 / 2];
#pragma unroll
            for (int n = 0; n < R2 / 2; ++n) {
                const int i = t + R1 * (e + 2 * n);
                v[n] = i < n_in ? b[i] : C2<T>{T(0), T(0)};
            }
            dft_pair<T, R2, false>(v, e);
#pragma unroll
            for (int k = 0; k < R2 / 2; ++k) {
                const int k2 = k + e * (R2 / 2);
                const C2<T> w = TL ? tw4[k2 * R1 + t] : ldg(tw4 + tw_fwd_gidx<R1, R2>(t, k2));
                tt[k2 * F::P1 + t] = k2 ? cmul(v[k], w) : v[k];
            }
        }
        __syncthreads();
        if (on && t < R2) {
            C2<T> u[R1 / 2];
#pragma unroll
            for (int n = 0; n < R1 / 2; ++n) u[n] = tt[t * F::P1 + e + 2 * n];
            dft_pair<T, R1, false>(u, e);
#pragma unroll
            for (int k = 0; k < R1 / 2; ++k) b[t + R2 * (k + e * (R1 / 2))] = u[k];
        }
        __syncthreads();
    }
}

// Inverse of fs_forward (unnormalised); only outputs n < n_out are written.
template <class T, int R1, int R2, bool TL = false>
__device__ void fs_inverse(C2<T> *buf, C2<T> *tb, const C2<T> *__restrict__ tw4, int nseq, int n_out, int sb) {
    if constexpr (sizeof(T) == 8) {
        fs_inverse1<T, R1, R2, TL>(buf, tb, tw4, nseq, n_out, sb);
        return;
    }
    using F = FourStep<R1, R2>;
    constexpr int L = 2 * F::TS;
    const int spb = blockDim.x / L;
    for (int base = 0; base < nseq; base += spb) {
        const int sq = base + threadIdx.x / L, r = threadIdx.x % L, t = r >> 1, e = r & 1;
        const bool on = sq < nseq && (int)threadIdx.x < spb * L;
        C2<T> *b = buf + sq * sb;
        C2<T> *tt = tb + (sq - base) * F::TB;
        if (on && t < R2) {
            C2<T> u[R1 / 2];
#pragma unroll
            for (int n = 0; n < R1 / 2; ++n) u[n] = b[t + R2 * (e + 2 * n)];
            dft_pair<T, R1, true>(u, e);
#pragma unroll
            for (int k = 0; k < R1 / 2; ++k) {
                const int t1 = k + e * (R1 / 2);
                C2<T> w = TL ? tw4[t * R1 + t1] : ldg(tw4 + (t * R1 + t1));
                w.y = -w.y;
                tt[t1 * F::P2 + t] = t1 ? cmul(u[k], w) : u[k];
            }
        }
        __syncthreads();
        if (on && t < R1) {
            C2<T> v[R2 / 2];
#pragma unroll
            for (int n = 0; n < R2 / 2; ++n) v[n] = tt[t * F::P2 + e + 2 * n];
            dft_pair<T, R2, true>(v, e);
#pragma unroll
            for (int k = 0; k < R2 / 2; ++k) {
                const int i = t + R1 * (k + e * (R2 / 2));
                if (i < n_out) b[i] = v[k];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------- three-step length-1024 FFTs --
// Length 1024 with 512 nonzero inputs / 512 kept outputs as three steps of
// 8-point register DFTs over 128 lanes: n = b + 64 m (b = t1 + 8 t1',
// t1, t1' < 8, m < 16), k = k2 + 128 k1 (k2 = k2' + 16 k1' < 128, k1 < 8).
// Forward: X[k2 + 128 k1] = sum_t1 W_8^{t1 k1} W_1024^{t1 k2} sum_t1'
// W_8^{t1' k1'} W_128^{t1' k2'} sum_m x[b + 64 m] W_16^{m k2'}; step A (lane
// pair b, e = lane & 1) takes the 16-point DFT over the 8 nonzero m as two
// 8-point DFTs, lane e the outputs k2' = 2j + e of x[m] W_16^{m e}; step B
// (lane t1 = lane & 7, k2' = lane >> 3) the 8-point DFT over t1'; step C
// (lane k2) over t1.  The inverse runs the steps backwards; its 16-point
// step keeps m < 8 as A + W_16^{-m} B of the pair's two 8-point DFTs.

// Transpose layout of steps A/B (and A'/B'): 32 lanes of a half-wave hit
// 32 distinct bank pairs on both sides (tests/test_three_step_fft_cpu.py
// checks it for every step and lane, and emulates the whole transform).
__device__ __forceinline__ int f8_ta(int t1, int t1p, int k2p) {
    return 4 * t1 + (k2p & 3) + 32 * (k2p >> 2) + 130 * t1p;
}
constexpr int kF8Ta = 1040;    // extent of f8_ta (complex values)
constexpr int kF8Tb = 1056;    // steps B/C transpose region: k2 + 132 t1
constexpr int kF8Tbs = 132;

// The lane's twiddles from its six bases (Args::tw8, [q][lane]):
// wa[j] = W^{8 t1' (2j + e)} = a0 a1^j, wb[j] = W^{t1 (k2' + 16 j)} = b0 b1^j,
// wc[j] = W^{j k2} = c1^j, wd[j] = W^{8 j k2'} = d1^j (W = W_1024).
template <class T>
__device__ __forceinline__ void f8_twiddles(const C2<T> (&q)[6], C2<T> (&wa)[8], C2<T> (&wb)[8], C2<T> (&wc)[8],
                                            C2<T> (&wd)[8]) {
    wa[0] = q[0];
    wb[0] = q[2];
    wc[0] = {T(1), T(0)};
    wd[0] = {T(1), T(0)};
    wc[1] = q[4];
    wd[1] = q[5];
#pragma unroll
    for (int j = 1; j < 8; ++j) {
        wa[j] = cmul(wa[j - 1], q[1]);
        wb[j] = cmul(wb[j - 1], q[3]);
        if (j > 1) {
            wc[j] = cmul(wc[j - 1], q[4]);
            wd[j] = cmul(wd[j - 1], q[5]);
        }
    }
}

// Forward three-step FFT of `nseq` sequences of LDS `buf` (stride sb >=
// kF8Ta), in place, natural order, inputs n < 512 (the rest is taken as
// zero and not read), 128 lanes per sequence; the step A/B transpose
// reuses the sequence's own buffer, tb holds one kF8Tb region per sequence
// of a round.  bases: the [6][128] twiddle bases, in LDS (TL) or global
// memory.
template <class T, bool TL>
__device__ void fs8_forward(C2<T> *buf, C2<T> *tb, const C2<T> *__restrict__ bases, int nseq, int sb) {
    const int spb = blockDim.x / 128;
    const int l = threadIdx.x & 127;
    const int e = l & 1, b = l >> 1, at1 = b & 7, at1p = b >> 3, bt1 = l & 7, bk2p = l >> 3, k2 = l;
    C2<T> q[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) q[i] = TL ? bases[i * 128 + l] : ldg(bases + i * 128 + l);
    C2<T> wa[8], wb[8], wc[8], wd[8];
    f8_twiddles(q, wa, wb, wc, wd);
    for (int base = 0; base < nseq; base += spb) {
        const int sq = base + (int)threadIdx.x / 128;
        const bool on = sq < nseq && (int)threadIdx.x < spb * 128;
        C2<T> *B = buf + sq * sb;
        C2<T> *TB = tb + (sq - base) * kF8Tb;
        C2<T> v[8];
        if (on) {
#pragma unroll
            for (int m = 0; m < 8; ++m) v[m] = B[b + 64 * m];
#pragma unroll
            for (int m = 1; m < 8; ++m) {
                const C2<T> w16 = {T(TwHolder<16>::tw.c[m]), T(-TwHolder<16>::tw.s[m])};
                const C2<T> r = cmul(v[m], w16);
                v[m] = e ? r : v[m];
            }
            DftReg<T, 8, false>::run(v);
        }
        __syncthreads();  // every input read before the transpose overwrites them
        if (on) {
#pragma unroll
            for (int j = 0; j < 8; ++j) B[f8_ta(at1, at1p, 2 * j + e)] = cmul(v[j], wa[j]);
        }
        __syncthreads();
        if (on) {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = B[f8_ta(bt1, j, bk2p)];
            DftReg<T, 8, false>::run(v);
#pragma unroll
            for (int j = 0; j < 8; ++j) TB[bk2p + 16 * j + kF8Tbs * bt1] = cmul(v[j], wb[j]);
        }
        __syncthreads();
        if (on) {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = TB[k2 + kF8Tbs * j];
            DftReg<T, 8, false>::run(v);
#pragma unroll
            for (int j = 0; j < 8; ++j) B[k2 + 128 * j] = v[j];
        }
        __syncthreads();
    }
}

// Inverse of fs8_forward (unnormalised): all 1024 inputs in natural order,
// outputs n < 512 written in natural order.
template <class T, bool TL>
__device__ void fs8_inverse(C2<T> *buf, C2<T> *tb, const C2<T> *__restrict__ bases, int nseq, int sb) {
    const int spb = blockDim.x / 128;
    const int l = threadIdx.x & 127;
    const int e = l & 1, b = l >> 1, at1 = b & 7, at1p = b >> 3, bt1 = l & 7, bk2p = l >> 3, k2 = l;
    C2<T> q[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) q[i] = TL ? bases[i * 128 + l] : ldg(bases + i * 128 + l);
    C2<T> wa[8], wb[8], wc[8], wd[8];
    f8_twiddles(q, wa, wb, wc, wd);
    for (int base = 0; base < nseq; base += spb) {
        const int sq = base + (int)threadIdx.x / 128;
        const bool on = sq < nseq && (int)threadIdx.x < spb * 128;
        C2<T> *B = buf + sq * sb;
        C2<T> *TB = tb + (sq - base) * kF8Tb;
        C2<T> v[8];
        if (on) {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = B[k2 + 128 * j];
            DftReg<T, 8, true>::run(v);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const C2<T> w = {wc[j].x, -wc[j].y};
                TB[k2 + kF8Tbs * j] = j ? cmul(v[j], w) : v[j];
            }
        }
        __syncthreads();
        if (on) {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = TB[bk2p + 16 * j + kF8Tbs * bt1];
            DftReg<T, 8, true>::run(v);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const C2<T> w = {wd[j].x, -wd[j].y};
                B[f8_ta(bt1, j, bk2p)] = j ? cmul(v[j], w) : v[j];
            }
        }
        __syncthreads();
        C2<T> o[4];
        if (on) {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = B[f8_ta(at1, at1p, 2 * j + e)];
            DftReg<T, 8, true>::run(v);
#pragma unroll
            for (int m = 1; m < 8; ++m) {
                const C2<T> w16 = {T(TwHolder<16>::tw.c[m]), T(TwHolder<16>::tw.s[m])};
                const C2<T> r = cmul(v[m], w16);
                v[m] = e ? r : v[m];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const C2<T> a0 = v[2 * i], a1 = v[2 * i + 1];
                const C2<T> p0 = {pair_swap(a0.x), pair_swap(a0.y)};
                const C2<T> p1 = {pair_swap(a1.x), pair_swap(a1.y)};
                // even lane: x[2i] = own + partner's; odd lane: x[2i + 1]
                const C2<T> mine = e ? a1 : a0, theirs = e ? p1 : p0;
                o[i] = e ? cadd(theirs, mine) : cadd(mine, theirs);
            }
        }
        __syncthreads();  // every transpose value read before the outputs overwrite them
        if (on) {
#pragma unroll
            for (int i = 0; i < 4; ++i) B[b + 64 * (2 * i + e)] = o[i];
        }
        __syncthreads();
    }
}

// Rows per work-group of the FP32 film's x passes with their inputs
// prefetched into LDS (Args::xpf), and their threads: eight waves, two per
// SIMD of the CU that holds the work-group (its LDS fits one work-group per
// CU), so that one wave's LDS and memory waits overlap the other's work.
constexpr int kXpfRows = 2;
constexpr int kXpfThreads = 512;

// ------------------------------------------------------------- x passes --
// An x-pass work-group owns `rows_per_wg` consecutive x-rows of the grid
// (row = y + ny z) and every cell and half-spectrum element of those rows.

// i / n for 0 <= i < 2^22: a compile-time n (four-step shapes) folds to a
// multiply-shift; otherwise a float reciprocal with one correction each way.
template <bool CONST_N>
__device__ __forceinline__ int small_div(int i, int n, float rn) {
    if constexpr (CONST_N) {
        return i / n;
    } else {
        int q = (int)((float)i * rn);
        q -= q * n > i ? 1 : 0;
        q += (q + 1) * n <= i ? 1 : 0;
        return q;
    }
}

// Row table of one work-group, filled once per launch: first cell of local
// row rr (-1 past the grid), its y and z, and the spectrum offset of packed
// line l = 3 rr + c at kx = 0 (-1 past the grid).
struct XRows {
    int cell0[kMaxRowsWg];
    int y[kMaxRowsWg], z[kMaxRowsWg];
    int spec[3 * kMaxRowsWg];
};

template <class T>
__device__ __forceinline__ void x_rows(const Args<T> &A, int r0, XRows &xr) {
    const int rr = threadIdx.x;
    if (rr < A.rows_per_wg) {
        const int row = r0 + rr;
        const bool ok = row < (int)A.nrows;
        const int z = ok ? row / A.ny : 0, y = ok ? row - z * A.ny : 0;
        xr.cell0[rr] = ok ? row * A.nx : -1;
        xr.y[rr] = y;
        xr.z[rr] = z;
#pragma unroll
        for (int c = 0; c < 3; ++c) xr.spec[3 * rr + c] = ok ? ((c * A.nzs + z) * A.nkx) * A.ly + y : -1;
    }
}

// Stage combination y_s = normalise(m0 + dt*scale*sum_q a_q k_q) over the
// work-group's rows (accumulated in FP64 for both storage precisions: the
// Dormand-Prince rows have coefficients up to ~12 that cancel to O(1)),
// y_s stored for s >= 2, Ms*y packed two real x-lines per complex sequence
// into b0.  NK = k-vectors of the combination (0: stage 1, y = m0).  Cells
// are taken in chunks of CH per thread (cell i = base + j*blockDim + tid);
// the loads of a chunk are all issued before any is consumed.
template <class T, int NK, int CH = Chunk<T>::value>
struct PackIn {
    T mv[CH][3], kv[CH][NK > 0 ? NK : 1][3], msv[CH];
};

// Issues the loads of one chunk; the k-vector at row position `skip` (-1:
// none) is not loaded (the caller has it in registers).
template <class T, int R1, int R2, int NK, int CH = Chunk<T>::value>
__device__ __forceinline__ void x_pack_load(const Args<T> &A, const Method &M, int s, int p, const XRows &xr,
                                            int base, int skip, PackIn<T, NK, CH> &in) {
    constexpr bool FAST = R1 > 0;
    const int px = FAST ? R1 * R2 : A.px, nx = px >> 1;
    const int ncl = A.rows_per_wg * nx;
    const float rnx = 1.0f / (float)nx;
    const int N = (int)A.ncell, bd = blockDim.x;
    const gp<const T> __restrict__ m0 = G(A.m_all) + (size_t)p * 3 * N;
    const StageRow &R = M.row[NK > 0 ? s : 2];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const int i = base + j * bd + (int)threadIdx.x;
        const int ic = i < ncl ? i : ncl - 1;
        const int rr = small_div<FAST>(ic, nx, rnx), x = ic - rr * nx;
        const int c0 = xr.cell0[rr];
        const int cell = (c0 < 0 ? 0 : c0) + x;
#pragma unroll
        for (int c = 0; c < 3; ++c) in.mv[j][c] = m0[c * N + cell];
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            if (q == skip) continue;
            const gp<const T> kp = G(A.k_all) + (size_t)kslot(A, M, R.kidx[q], p) * 3 * N;
#pragma unroll
            for (int c = 0; c < 3; ++c) in.kv[j][q][c] = kp[c * N + cell];
        }
        in.msv[j] = A.ms_pack ? G(A.ms_pack)[cell] : A.ms;
    }
}

// Combination, normalisation, store and packing of one loaded chunk.
template <class T, int R1, int R2, int NK, int CH = Chunk<T>::value>
__device__ __forceinline__ void x_pack_put(const Args<T> &A, const Method &M, int s, int p, double dt,
                                           const XRows &xr, int base, const PackIn<T, NK, CH> &in, C2<T> *b0) {
    constexpr bool FAST = R1 > 0;
    const int px = FAST ? R1 * R2 : A.px, nx = px >> 1;
    const int ncl = A.rows_per_wg * nx;
    const float rnx = 1.0f / (float)nx;
    const int N = (int)A.ncell, bd = blockDim.x;
    const StageRow &R = M.row[NK > 0 ? s : 2];
    const double f = dt * R.scale;
    const int sb = FAST ? A.xsb : px;
    const gp<T> __restrict__ yout = NK > 0 ? stage_output(A, M, s, p) : nullptr;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const int i = base + j * bd + (int)threadIdx.x;
        if (i >= ncl) continue;
        const int rr = small_div<FAST>(i, nx, rnx), x = i - rr * nx;
        const int c0 = xr.cell0[rr];
        T v[3];
        if constexpr (NK == 0) {
            v[0] = in.mv[j][0];
            v[1] = in.mv[j][1];
            v[2] = in.mv[j][2];
            if (A.rec_y && c0 >= 0) {
                const int cell = c0 + x;
                T *ry = A.rec_y + rec_slot(A, A.rec_sv);
                ry[cell] = v[0];
                ry[N + cell] = v[1];
                ry[2 * N + cell] = v[2];
            }
        } else {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
            for (int q = 0; q < NK; ++q) {
                const double w = R.a[q];
                a0 += w * (double)in.kv[j][q][0];
                a1 += w * (double)in.kv[j][q][1];
                a2 += w * (double)in.kv[j][q][2];
            }
            if (A.errv && s == M.nstage && c0 >= 0) {
                // error terms of the final row (Method::efuse: the same k's,
                // in the same order, as the error estimate's)
                double e0 = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll
                for (int q = 0; q < NK; ++q) {
                    const double w = M.e[q];
                    e0 += w * (double)in.kv[j][q][0];
                    e1 += w * (double)in.kv[j][q][1];
                    e2 += w * (double)in.kv[j][q][2];
                }
                const int cell = c0 + x;
                G(A.errv)[cell] = e0;
                G(A.errv)[N + cell] = e1;
                G(A.errv)[2 * N + cell] = e2;
            }
            const double xx = (double)in.mv[j][0] + a0 * f, yy = (double)in.mv[j][1] + a1 * f,
                         zz = (double)in.mv[j][2] + a2 * f;
            const double n2 = xx * xx + yy * yy + zz * zz;
            const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
            v[0] = (T)(xx * inv);
            v[1] = (T)(yy * inv);
            v[2] = (T)(zz * inv);
            if (c0 >= 0) {
                const int cell = c0 + x;
                yout[cell] = v[0];
                yout[N + cell] = v[1];
                yout[2 * N + cell] = v[2];
                if (A.rec_y) {
                    T *ry = A.rec_y + rec_y_off(A, M, s);
                    ry[cell] = v[0];
                    ry[N + cell] = v[1];
                    ry[2 * N + cell] = v[2];
                    G(A.rec_vn)[rec_slot(A, A.rec_ss) + (size_t)(s - 2) * N + cell] = (T)sqrt(n2);
                }
            }
        }
        const T w = c0 >= 0 ? in.msv[j] : T(0);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int l = 3 * rr + c;
            T *e = reinterpret_cast<T *>(&b0[(l >> 1) * sb + x]);
            e[l & 1] = w * v[c];
        }
    }
}

// Zero imaginary half of the last sequence when the line count is odd.
template <class T, int R1, int R2>
__device__ __forceinline__ void x_pack_tail(const Args<T> &A, C2<T> *b0) {
    const int px = R1 > 0 ? R1 * R2 : A.px, nx = px >> 1, nl = 3 * A.rows_per_wg;
    const int sb = R1 > 0 ? A.xsb : px;
    if (nl & 1)
        for (int x = threadIdx.x; x < nx; x += blockDim.x) b0[(nl >> 1) * sb + x].y = T(0);
}

template <class T, int R1, int R2, int NK, int CH = Chunk<T>::value>
__device__ __forceinline__ void x_pack_stage(const Args<T> &A, const Method &M, int s, int p, double dt,
                                             const XRows &xr, C2<T> *b0) {
    const int px = R1 > 0 ? R1 * R2 : A.px, ncl = A.rows_per_wg * (px >> 1);
    for (int base = 0; base < ncl; base += CH * (int)blockDim.x) {
        PackIn<T, NK, CH> in;
        x_pack_load<T, R1, R2, NK, CH>(A, M, s, p, xr, base, -1, in);
        x_pack_put<T, R1, R2, NK, CH>(A, M, s, p, dt, xr, base, in, b0);
    }
    x_pack_tail<T, R1, R2>(A, b0);
}

// Zero-padded forward FFT of the packed sequences in b0 and the unpacking of
// each sequence into the half spectra of its two real lines,
// X = (Z[k] + conj Z[-k])/2 and (Z[k] - conj Z[-k])/(2i), y-lines contiguous.
// F8: the three-step FFT (fs8_forward; R1 = R2 = 32, FP32), `tw` then
// holding its twiddle bases when TL.
template <class T, int R1, int R2, bool TL = false, bool F8 = false>
__device__ __forceinline__ void x_forward_store(const Args<T> &A, const XRows &xr, C2<T> *b0, C2<T> *b1,
                                                const C2<T> *tw) {
    constexpr bool FAST = R1 > 0;
    const int px = FAST ? R1 * R2 : A.px, nx = px >> 1, nkx = nx + 1;
    __syncthreads();
    const C2<T> *res;
    const int sb = FAST ? A.xsb : px;
    if constexpr (FAST) {
        // TL: `tw` holds the four-step twiddles in LDS
        if constexpr (F8)
            fs8_forward<T, TL>(b0, b1, TL ? tw : reinterpret_cast<const C2<T> *>(A.tw8), A.nseq_x, sb);
        else
            fs_forward<T, R1, R2, TL>(b0, b1, TL ? tw : reinterpret_cast<const C2<T> *>(A.tw4x), A.nseq_x, nx, sb);
        res = b0;
    } else {
        res = lds_fft<T, false>(b0, b1, A.fx, tw, A.nseq_x, nx, px);
    }
    // lanes take the lines of one kx together (line fastest): the rows of a
    // component are consecutive y-points, so one store instruction writes
    // runs of rows_per_wg values instead of single values a y-line apart
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec);
    const int ly = A.ly, nl = 3 * A.rows_per_wg, tot = nl * nkx;
    const float rnl = 1.0f / (float)nl;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int kx = small_div<false>(i, nl, rnl), l = i - kx * nl;
        const int off = xr.spec[l];
        if (off < 0) continue;
        const C2<T> *zq = res + (l >> 1) * sb;
        const C2<T> zk = zq[kx], zm = zq[kx == 0 ? 0 : px - kx];
        const C2<T> X = (l & 1) ? C2<T>{(zk.y + zm.y) * T(0.5), (zm.x - zk.x) * T(0.5)}
                                : C2<T>{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
        stg(S + off + kx * ly, X);
    }
}

// Hermitian rebuild of the packed sequences from the half spectra (two real
// output lines per complex sequence) and the C2R inverse FFT along x, only
// the first nx outputs kept.  Lanes take the lines of one kx together (line
// fastest, the two lines of a packed sequence in an even/odd lane pair, so
// that the rows of a component are consecutive y-points of one load); each
// half-spectrum value is loaded once, and the even lane writes Z = A + iB at
// kx and, from the conjugates, at px - kx.
template <class T, int R1, int R2, bool F8 = false>
__device__ __forceinline__ const C2<T> *x_inverse_load(const Args<T> &A, const XRows &xr, C2<T> *b0, C2<T> *b1,
                                                       const C2<T> *tw) {
    constexpr bool FAST = R1 > 0;
    constexpr int CZ = 8;  // loads in flight per thread
    const int px = FAST ? R1 * R2 : A.px, nx = px >> 1, nkx = nx + 1;
    const int sb = FAST ? A.xsb : px;
    const int nlp = 2 * A.nseq_x, nl = 3 * A.rows_per_wg, ly = A.ly, bd = blockDim.x;
    const int tot = nlp * nkx;  // even: lane pairs never straddle its end
    const float rnl = 1.0f / (float)nlp;
    const C2<T> *S = reinterpret_cast<const C2<T> *>(A.spec);
    for (int base = 0; base < tot; base += CZ * bd) {
        C2<T> v[CZ];
#pragma unroll
        for (int j = 0; j < CZ; ++j) {
            const int i = base + j * bd + (int)threadIdx.x;
            const int ic = i < tot ? i : tot - 1;
            const int kk = small_div<false>(ic, nlp, rnl), l = ic - kk * nlp;
            const int off = l < nl ? xr.spec[l] : -1;
            // unconditional load (a conditional one becomes a branch per
            // load), zero for lines past the grid
            const C2<T> w = ldg(S + (off >= 0 ? off + kk * ly : 0));
            v[j] = off >= 0 ? w : C2<T>{T(0), T(0)};
        }
#pragma unroll
        for (int j = 0; j < CZ; ++j) {
            const int i = base + j * bd + (int)threadIdx.x;
            // every lane takes part in the exchange (wave-uniform control flow)
            const C2<T> b = {pair_swap(v[j].x), pair_swap(v[j].y)};
            if (i >= tot || (threadIdx.x & 1)) continue;
            const int kk = small_div<false>(i, nlp, rnl), l = i - kk * nlp;
            const C2<T> a = v[j];
            C2<T> *q = b0 + (l >> 1) * sb;
            q[kk] = {a.x - b.y, a.y + b.x};
            if (kk > 0 && kk < nx) q[px - kk] = {a.x + b.y, b.x - a.y};
        }
    }
    __syncthreads();
    if constexpr (FAST) {
        if constexpr (F8)
            fs8_inverse<T, false>(b0, b1, reinterpret_cast<const C2<T> *>(A.tw8), A.nseq_x, sb);
        else
            fs_inverse<T, R1, R2>(b0, b1, reinterpret_cast<const C2<T> *>(A.tw4x), A.nseq_x, nx, sb);
        return b0;
    } else {
        return lds_fft<T, true>(b0, b1, A.fx, tw, A.nseq_x, px, nx);
    }
}

// Local pass of stage s over the work-group's rows: demag field from the
// inverse x FFT (1/P scale), exchange, Zeeman, anisotropy, LLG torque,
// optional statistics and error norm.  x_local_load issues the loads of one
// chunk of cells, x_local_put consumes it (and returns the chunk's k).
template <class T, int FT, int R1, int R2, int CH = Chunk<T>::value>
__device__ __forceinline__ void x_local_load(const Args<T> &A, const Method &M, int p, gp<const T> yin,
                                             const XRows &xr, int base, bool err, CellIn<T> *g) {
    constexpr bool FAST = R1 > 0;
    const int px = FAST ? R1 * R2 : A.px, nx = px >> 1;
    const int ncl = A.rows_per_wg * nx, bd = blockDim.x;
    const float rnx = 1.0f / (float)nx;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const int i = base + j * bd + (int)threadIdx.x;
        const int ic = i < ncl ? i : ncl - 1;
        const int rr = small_div<FAST>(ic, nx, rnx), x = ic - rr * nx;
        const int c0 = xr.cell0[rr];
        gather_cell<T, FT>(A, M, p, (c0 < 0 ? 0 : c0) + x, x, xr.y[rr], xr.z[rr], yin, err, g[j]);
    }
}

template <class T, int FT, int R1, int R2, int CH = Chunk<T>::value>
__device__ __forceinline__ void x_local_put(const Args<T> &A, const Method &M, int s, int p, double dt, TimeF<T> tf,
                                            const XRows &xr, int base, const CellIn<T> *g, const C2<T> *res,
                                            bool stats, bool err, Acc &acc, T (*kk)[3]) {
    constexpr bool FAST = R1 > 0;
    const int px = FAST ? R1 * R2 : A.px, nx = px >> 1;
    const int ncl = A.rows_per_wg * nx, bd = blockDim.x;
    const float rnx = 1.0f / (float)nx;
    const gp<T> kout = stage_rhs(A, M, s, p);
    const gp<T> kpush = stage_push(A, M, s);
    const gp<T> hrec = rec_h_ptr(A, M, s);
    const T scale = T(A.inv_p);
    const int sb = FAST ? A.xsb : px;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const int i = base + j * bd + (int)threadIdx.x;
        if (i >= ncl) continue;
        const int rr = small_div<FAST>(i, nx, rnx), x = i - rr * nx;
        const int c0 = xr.cell0[rr];
        if (c0 < 0) continue;
        T hd[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int l = 3 * rr + c;
            const C2<T> zv = res[(l >> 1) * sb + x];
            hd[c] = ((l & 1) ? zv.y : zv.x) * scale;
        }
        compute_cell<T, FT>(A, M, c0 + x, dt, g[j], kout, hd, tf, stats, err, acc, hrec, kk ? kk[j] : nullptr,
                             kpush);
    }
}

template <class T, int FT, int R1, int R2, int CH = Chunk<T>::value>
__device__ __forceinline__ void x_local(const Args<T> &A, const Method &M, int s, int p, double dt, TimeF<T> tf,
                                        const XRows &xr, const C2<T> *res, bool stats, bool err, Acc &acc) {
    const int px = R1 > 0 ? R1 * R2 : A.px, ncl = A.rows_per_wg * (px >> 1);
    const gp<const T> yin = stage_input(A, M, s, p);
    for (int base = 0; base < ncl; base += CH * (int)blockDim.x) {
        CellIn<T> g[CH];
        x_local_load<T, FT, R1, R2, CH>(A, M, p, yin, xr, base, err, g);
        x_local_put<T, FT, R1, R2, CH>(A, M, s, p, dt, tf, xr, base, g, res, stats, err, acc, nullptr);
    }
}

// Forward x pass of stage s: combination, Ms*y packing, R2C FFT along x.
// TPB: work-group size (kThreads, or kXpfThreads for the FP32 film, whose
// work-groups are one per CU: two waves per SIMD); CH: cells per thread of
// a chunk.
template <class T, int R1, int R2, int NK, int TPB = kThreads, int CH = Chunk<T>::value>
__global__ void __launch_bounds__(TPB) k_fwd_x(const Args<T> *__restrict__ AP, int s) {
    const Args<T> &A = AP[blockIdx.y];
    const CtlR ctl = ctl_ro(A);
    // the control fields in one round trip, then the early exit
    const int done = ctl->done, p = ctl->parity;
    const double dt = ctl->dt;
    if (done) return;
    const Method &M = method_for(A);
    C2<T> *b0 = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *b1 = b0 + A.nseq_x * A.xsb;
    C2<T> *tw = b1 + A.nseq_x * A.xsb;
    __shared__ XRows xr;
    x_rows(A, (int)xcd_block(blockIdx.x, gridDim.x) * A.rows_per_wg, xr);
    if constexpr (R1 == 0) load_twiddles<T>(tw, A.twx, A.px);
    __syncthreads();
    x_pack_stage<T, R1, R2, NK, CH>(A, M, s, p, dt, xr, b0);
    // the FP32 film's px = 1024 with kXpfThreads: three-step FFT
    constexpr bool F8 = sizeof(T) == 4 && R1 == 32 && R2 == 32 && TPB == kXpfThreads;
    x_forward_store<T, R1, R2, false, F8>(A, xr, b0, b1, tw);
}

// Inverse x pass and local pass of stage s, with the work-group partials of
// the statistics / error norm when asked (k_control reduces them).  TPB and
// CH as in k_fwd_x.
template <class T, int FT, int R1, int R2, int TPB = kThreads, int CH = Chunk<T>::value>
__global__ void __launch_bounds__(TPB) k_inv_x(const Args<T> *__restrict__ AP, int s, int stats) {
    const Args<T> &A = AP[blockIdx.y];
    const CtlR ctl = ctl_ro(A);
    // the control fields in one round trip, then the early exit
    const int done = ctl->done, p = ctl->parity;
    const double dt = ctl->dt;
    if (done) return;
    const Method &M = method_for(A);
    C2<T> *b0 = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *b1 = b0 + A.nseq_x * A.xsb;
    C2<T> *tw = b1 + A.nseq_x * A.xsb;
    __shared__ XRows xr;
    x_rows(A, (int)xcd_block(blockIdx.x, gridDim.x) * A.rows_per_wg, xr);
    if constexpr (R1 == 0) load_twiddles<T>(tw, A.twx, A.px);
    __syncthreads();
    constexpr bool F8 = sizeof(T) == 4 && R1 == 32 && R2 == 32 && TPB == kXpfThreads;
    const C2<T> *res = x_inverse_load<T, R1, R2, F8>(A, xr, b0, b1, tw);
    const bool err = M.adaptive && s == M.nstage;
    const TimeF<T> tf = time_factors(A, stage_time(A, M, s, ctl, dt));
    Acc acc;
    acc.init();
    x_local<T, FT, R1, R2, CH>(A, M, s, p, dt, tf, xr, res, stats != 0, err, acc);
    if (stats || err) block_reduce_global(acc, A.partials + (long long)blockIdx.x * kStatSlots);
}

// Waves per SIMD the joined x pass is compiled for: FP32 with the uniform
// exchange, no further terms and px <= 1024 fits 168 VGPRs (three
// work-groups per CU, measured 11% faster on a 256x256x64 block than two);
// FP64, the region table, the further terms and px = 2048 would spill.
template <class T, int FT, int R1, int R2>
constexpr int xx_waves() {
    return sizeof(T) == 4 && FT == 0 && R1 * R2 <= 1024 ? 3 : 1;
}

// Stage s (not the last of the step) ends and stage s+1 begins in one
// launch: the combination of stage s+1 is pointwise in k_1..k_s, and the
// work-group owns all cells of its rows, so the inverse x pass, local pass
// and k_s of stage s feed the forward x pass of stage s+1 directly.  NK =
// k-vectors of stage s+1.  Results are those of the separate launches.
// (Issuing all three phases' loads up front was measured slower: the
// register footprint doubles and the 3-D grid loses two thirds of its
// occupancy.)
template <class T, int FT, int R1, int R2, int NK>
__global__ void __launch_bounds__(kThreads, (xx_waves<T, FT, R1, R2>())) k_xx(const Args<T> *__restrict__ AP, int s) {
    const Args<T> &A = AP[blockIdx.y];
    const CtlR ctl = ctl_ro(A);
    // the control fields in one round trip, then the early exit
    const int done = ctl->done, p = ctl->parity;
    const double dt = ctl->dt;
    if (done) return;
    const Method &M = method_for(A);
    C2<T> *b0 = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *b1 = b0 + A.nseq_x * A.xsb;
    C2<T> *tw = b1 + A.nseq_x * A.xsb;
    __shared__ XRows xr;
    x_rows(A, (int)xcd_block(blockIdx.x, gridDim.x) * A.rows_per_wg, xr);
    if constexpr (R1 == 0) load_twiddles<T>(tw, A.twx, A.px);
    __syncthreads();
    const C2<T> *res = x_inverse_load<T, R1, R2>(A, xr, b0, b1, tw);
    const TimeF<T> tf = time_factors(A, stage_time(A, M, s, ctl, dt));
    Acc acc;
    x_local<T, FT, R1, R2>(A, M, s, p, dt, tf, xr, res, false, false, acc);
    // k_s of these rows is written by this work-group; the barrier publishes
    // it (work-group scope) and frees the LDS result buffer
    __syncthreads();
    x_pack_stage<T, R1, R2, NK>(A, M, s + 1, p, dt, xr, b0);
    x_forward_store<T, R1, R2>(A, xr, b0, b1, tw);
}

// ---------------------------------------------- film joined x pass, LDS prefetch --
// On the 512 x 512 film every launch is one round of work-groups (one per
// CU, one wave per SIMD), so k_xx lasts one work-group's dependent chain:
// spectrum loads -> inverse FFT -> stencil loads -> local pass -> stage
// combination loads -> forward FFT -> stores.  k_xxp issues every global
// read of the launch at its start: the spectrum into registers, and the
// stencil rows, the activity bytes and the stage combination's m0 and k
// vectors straight into LDS with global->LDS loads (no VGPRs held), so the
// three round trips overlap.  The values, and hence the arithmetic and the
// results, are those of k_xx.  Film (nz = 1), two rows per work-group,
// nx a multiple of 256, uniform exchange without further terms, one-step
// methods.
//
// LDS after the FFT buffers (floats unless noted):
//   Y  [4][3][nx]  stencil rows y0-1 .. y0+2 (clamped to the grid) of y_s
//   M  [3][2][nx]  m0 of the two rows
//   K  [NK][3][2][nx] the k-vectors of stage s+1's combination (the slot of
//                  k_s is filled by the local pass)
//   TW [2nx] complex  the four-step twiddles of the x FFTs
//   NM [2][nx] bytes  activity bytes of the two rows (with a mask)

// The FP32 film's x passes (Args::xpf: one work-group per CU) run
// kXpfThreads threads per work-group, xpf_chunk cells per thread of a chunk.
template <int R1, int R2>
constexpr int xpf_chunk() {
    return kXpfRows * (R1 * R2 / 2) > kXpfThreads ? kXpfRows * (R1 * R2 / 2) / kXpfThreads : 1;
}
template <class T, int R1, int R2>
constexpr bool xpf_shape() {
    return sizeof(T) == 4 && (R1 * R2 == 512 || R1 * R2 == 1024);
}

template <class T>
__host__ __device__ constexpr size_t xpf_floats(int nx, int nk) {
    return (size_t)(4 * 3 + 3 * kXpfRows + nk * 3 * kXpfRows + 4) * nx;
}

// Moves n floats (n a multiple of 256; src 16-byte aligned) from global src
// to LDS dst with direct global->LDS loads of the calling wave: lane l of
// load j moves floats 256 j + 4 l .. + 3.  dst must be wave-uniform and all
// 64 lanes of the wave active (an absent lane leaves its 16 bytes unset).
__device__ __forceinline__ void dma_floats(const float *src, float *dst, int n) {
    const int lane = threadIdx.x & 63;
    for (int j = 0; j < n; j += 256)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + j + 4 * lane),
                                         (__attribute__((address_space(3))) void *)(dst + j), 16, 0, 0);
}
// The same for n bytes (n a multiple of 1024).
__device__ __forceinline__ void dma_bytes(const uint8_t *src, uint8_t *dst, int n) {
    const int lane = threadIdx.x & 63;
    for (int j = 0; j < n; j += 1024)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + j + 16 * lane),
                                         (__attribute__((address_space(3))) void *)(dst + j), 16, 0, 0);
}

// CellIn of a film cell from the prefetched rows (the loads gather_cell<T,0>
// makes, with the film's neighbour rule: x and y clamped to the grid, z
// neighbours the cell itself).
template <class T>
__device__ __forceinline__ void gather_film_lds(const Args<T> &A, int cell, int x, int rr, const T *Y,
                                                const uint8_t *NM, CellIn<T> &g) {
    const int nx = A.nx, N = (int)A.ncell;
    const unsigned bits = A.nmask ? (unsigned)NM[rr * nx + x] : 0x7Fu;
    g.act = bits & 1u;
#pragma unroll
    for (int q = 0; q < 6; ++q) g.nact[q] = (bits >> (q + 1)) & 1u;
    const int xm = x > 0 ? x - 1 : x, xp = x < nx - 1 ? x + 1 : x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T *r0 = Y + (rr * 3 + c) * nx, *r1 = r0 + 3 * nx, *r2 = r1 + 3 * nx;
        g.m[c] = r1[x];
        g.v[0][c] = r1[xm];
        g.v[1][c] = r1[xp];
        g.v[2][c] = r0[x];
        g.v[3][c] = r2[x];
        g.v[4][c] = r1[x];
        g.v[5][c] = r1[x];
    }
    g.ms = A.ms_cell ? G(A.ms_cell)[cell] : A.ms;
    if (A.has_uni) {
        g.ku1 = A.ku1_cell ? G(A.ku1_cell)[cell] : A.ku1_u;
        g.ku2 = A.ku2_cell ? G(A.ku2_cell)[cell] : A.ku2_u;
#pragma unroll
        for (int c = 0; c < 3; ++c) g.ax[c] = A.axis_cell ? G(A.axis_cell)[c * N + cell] : A.axis[c];
    }
    g.wm = A.wave_mask ? G(A.wave_mask)[cell] : 1;
}

template <class T, int R1, int R2, int NK>
__global__ void __launch_bounds__(kXpfThreads) k_xxp(const Args<T> *__restrict__ AP, int s) {
    static_assert(sizeof(T) == 4, "FP32 only");
    constexpr int PX = R1 * R2, NX = PX / 2, NKX = NX + 1;
    // cells per thread of the local pass and the combination
    constexpr int CH = kXpfRows * NX > kXpfThreads ? kXpfRows * NX / kXpfThreads : 1;
    constexpr int NLP = 2 * ((3 * kXpfRows + 1) / 2);           // packed lines (lane pairs)
    constexpr int TOT = NLP * NKX;                              // spectrum values of the work-group
    constexpr int NV = (TOT + kXpfThreads - 1) / kXpfThreads;   // per thread
    const Args<T> &A = AP[blockIdx.y];
    const int N = (int)A.ncell, sb = A.xsb, ly = A.ly, nkx = A.nkx;
    C2<T> *b0 = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *b1 = b0 + A.nseq_x * A.xsb;
    T *Y = reinterpret_cast<T *>(fm_smem + A.xpf_off);
    T *Mz = Y + 12 * NX;
    T *K = Mz + 3 * kXpfRows * NX;
    T *TW = K + NK * 3 * kXpfRows * NX;
    uint8_t *NM = reinterpret_cast<uint8_t *>(TW + 2 * PX);
    __shared__ XRows xr;
    const int r0 = (int)xcd_block(blockIdx.x, gridDim.x) * kXpfRows;
    // the work-group's rows are y0, y0 + 1 of the film (nrows even)
    const int y0 = r0, c0 = r0 * NX;
    // 1. the loads that do not depend on the control block go out first, so
    // that their latency overlaps the control block's round trip: the
    // spectrum values into registers (as x_inverse_load: line fastest), the
    // twiddles and the activity bytes into LDS (the work-group's loads are
    // wasted when the device loop has finished: the early exit follows)
    const C2<T> *S = reinterpret_cast<const C2<T> *>(A.spec);
    constexpr int NL = 3 * kXpfRows;
    const float rnl = 1.0f / (float)NLP;
    C2<T> v[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const int i = j * kXpfThreads + (int)threadIdx.x;
        const int ic = i < TOT ? i : TOT - 1;
        const int kk = small_div<false>(ic, NLP, rnl), l = ic - kk * NLP;
        const int rr = l / 3, c = l - 3 * rr;  // line l = 3 rr + c, as XRows::spec
        static_assert(NL == NLP, "two rows pack into whole sequences");
        // unconditional loads: a conditional one becomes a branch per load,
        // and the compiler then re-reads the arguments inside each branch
        v[j] = ldg(S + (c * nkx) * ly + y0 + rr + kk * ly);
    }
    const int wv = threadIdx.x >> 6, nw = kXpfThreads / 64;
    // px = 1024: three-step FFTs, TW holds their [6][128] twiddle bases
    constexpr bool F8 = R1 == 32 && R2 == 32;
    if (wv == nw - 1) {
        if constexpr (F8) dma_floats(reinterpret_cast<const T *>(A.tw8), TW, 2 * 6 * 128);
        else dma_floats(reinterpret_cast<const T *>(A.tw4x), TW, 2 * PX);
    }
    if (A.nmask && wv == nw - 2) dma_bytes(A.nmask + c0, NM, kXpfRows * NX);
    // the control block through the scalar cache (this kernel does not
    // write it; the scalar cache is invalidated at every launch): its round
    // trip counts on the scalar counter and does not wait behind the
    // spectrum loads, which count on the vector counter in issue order
    const CtlR ctl = ctl_ro(A);
    const int done = ctl->done, p = ctl->parity;
    const double dt = ctl->dt;
    if (done) {
        // the global->LDS loads land before the work-group's LDS is released
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        return;
    }
    const Method &M = method_for(A);
    x_rows(A, r0, xr);
    // stage s+1's row: which of its k-vectors is k_s (made here), and the
    // source addresses of the remaining prefetch
    const StageRow &R = M.row[s + 1];
    int qs = -1;
    const T *ksrc[NK];
#pragma unroll
    for (int q = 0; q < NK; ++q) {
        if (R.kidx[q] == s) qs = q;
        ksrc[q] = A.k_all + (size_t)kslot(A, M, R.kidx[q], p) * 3 * N + c0;
    }
    const T *ysrc = (const T *)stage_input(A, M, s, p);
    const T *msrc = A.m_all + (size_t)p * 3 * N + c0;
    // 2. the rest of the launch's inputs into LDS, spread over the waves
    // (the twiddle wave takes no rows: its only global->LDS loads are the
    // twiddles, so that the inverse FFT waits for those alone and runs
    // while the rows, needed from the local pass on, still arrive)
    {
        const int nwr = nw - 1;
        int item = 0;
        for (int k = 0; k < 4; ++k) {
            const int row = min(max(y0 - 1 + k, 0), A.ny - 1);
            for (int c = 0; c < 3; ++c, ++item)
                if (item % nwr == wv) dma_floats(ysrc + c * N + row * NX, Y + (k * 3 + c) * NX, NX);
        }
        for (int c = 0; c < 3; ++c, ++item)
            if (item % nwr == wv) dma_floats(msrc + c * N, Mz + c * kXpfRows * NX, kXpfRows * NX);
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            if (q == qs) continue;
            for (int c = 0; c < 3; ++c, ++item)
                if (item % nwr == wv) dma_floats(ksrc[q] + c * N, K + (q * 3 + c) * kXpfRows * NX, kXpfRows * NX);
        }
    }
    // 3. Hermitian rebuild of the packed sequences, inverse FFT
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const int i = j * kXpfThreads + (int)threadIdx.x;
        const C2<T> b = {pair_swap(v[j].x), pair_swap(v[j].y)};
        if (i >= TOT || (threadIdx.x & 1)) continue;
        const int kk = small_div<false>(i, NLP, rnl), l = i - kk * NLP;
        const C2<T> a = v[j];
        C2<T> *q = b0 + (l >> 1) * sb;
        q[kk] = {a.x - b.y, a.y + b.x};
        if (kk > 0 && kk < NX) q[PX - kk] = {a.x + b.y, b.x - a.y};
    }
    // the twiddles have arrived (the twiddle wave waits for its
    // global->LDS loads; the barrier publishes them)
    if (wv == nw - 1) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const C2<T> *tw = reinterpret_cast<const C2<T> *>(TW);
    if constexpr (F8) fs8_inverse<T, true>(b0, b1, tw, A.nseq_x, sb);
    else fs_inverse<T, R1, R2, true>(b0, b1, tw, A.nseq_x, NX, sb);
    // the prefetched rows have arrived (every wave waits for its own
    // global->LDS loads, the barrier publishes them)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    // 5. local pass of stage s from the LDS rows; k_s also into its K slot
    const TimeF<T> tf = time_factors(A, stage_time(A, M, s, ctl, dt));
    {
        Acc acc;
        const gp<T> kout = stage_rhs(A, M, s, p);
        const gp<T> hrec = rec_h_ptr(A, M, s);
        const T scale = T(A.inv_p);
        constexpr int NCL = kXpfRows * NX;
        for (int base = 0; base < NCL; base += CH * kXpfThreads) {
            CellIn<T> g[CH];
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int i = base + j * kXpfThreads + (int)threadIdx.x;
                const int ic = i < NCL ? i : NCL - 1;
                const int rr = ic / NX, x = ic - rr * NX;
                gather_film_lds(A, c0 + ic, x, rr, Y, NM, g[j]);
            }
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int i = base + j * kXpfThreads + (int)threadIdx.x;
                if (i >= NCL) continue;
                const int rr = i / NX, x = i - rr * NX;
                T hd[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int l = 3 * rr + c;
                    const C2<T> zv = b0[(l >> 1) * sb + x];
                    hd[c] = ((l & 1) ? zv.y : zv.x) * scale;
                }
                T kr[3];
                compute_cell<T, 0>(A, M, c0 + i, dt, g[j], kout, hd, tf, false, false, acc, hrec, kr, nullptr);
                if (qs >= 0)
#pragma unroll
                    for (int c = 0; c < 3; ++c) K[(qs * 3 + c) * kXpfRows * NX + i] = kr[c];
            }
        }
    }
    __syncthreads();
    // 6. combination of stage s+1 from LDS, Ms*y packed, forward FFT, store
    {
        constexpr int NCL = kXpfRows * NX;
        for (int base = 0; base < NCL; base += CH * kXpfThreads) {
            PackIn<T, NK, CH> in;
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int i = base + j * kXpfThreads + (int)threadIdx.x;
                const int ic = i < NCL ? i : NCL - 1;
#pragma unroll
                for (int c = 0; c < 3; ++c) in.mv[j][c] = Mz[c * NCL + ic];
#pragma unroll
                for (int q = 0; q < NK; ++q)
#pragma unroll
                    for (int c = 0; c < 3; ++c) in.kv[j][q][c] = K[(q * 3 + c) * NCL + ic];
                in.msv[j] = A.ms_pack ? G(A.ms_pack)[c0 + ic] : A.ms;
            }
            x_pack_put<T, R1, R2, NK, CH>(A, M, s + 1, p, dt, xr, base, in, b0);
        }
        x_pack_tail<T, R1, R2>(A, b0);
    }
    x_forward_store<T, R1, R2, true, F8>(A, xr, b0, b1, tw);
}

// MODE 0: thin-film fused forward FFT + tensor multiply + inverse FFT.
// MODE 1: 3-D forward y FFT.   MODE 2: 3-D inverse y FFT.
template <class T, int MODE>
__global__ void __launch_bounds__(kThreads) k_y(const Args<T> *__restrict__ AP) {
    const Args<T> &A = AP[blockIdx.y];
    if (ctl_ro(A)->done) return;
    const int py = A.py, ny = A.ny, ly = A.ly, nkx = A.nkx;
    C2<T> *b0 = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *b1 = b0 + 3 * py;
    C2<T> *tw = b1 + 3 * py;
    const int kx = MODE == 0 ? blockIdx.x : blockIdx.x % nkx;
    const int zz = MODE == 0 ? 0 : blockIdx.x / nkx;
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec);
    load_twiddles<T>(tw, A.twy, py);
    const int nload = MODE == 2 ? py : ny;
    for (int i = threadIdx.x; i < 3 * nload; i += blockDim.x) {
        const int c = i / nload, yy = i - c * nload;
        b0[c * py + yy] = S[(((long long)c * A.nzs + zz) * nkx + kx) * ly + yy];
    }
    __syncthreads();
    C2<T> *res;
    int nstore;
    if (MODE == 0) {
        res = lds_fft<T, false>(b0, b1, A.fy, tw, 3, ny, py);
        const long long cs = (long long)nkx * (ny + 1);
        const T *N = A.ntens + (long long)kx * (ny + 1);
        if (A.gtens) {
            // complex spectrum of a periodic kernel: [c][kx][ky], all ky
            const C2<T> *G = reinterpret_cast<const C2<T> *>(A.gtens) + (long long)kx * py;
            const long long gs = (long long)nkx * py;
            for (int ky = threadIdx.x; ky < py; ky += blockDim.x) {
                const C2<T> gxx = G[ky], gyy = G[gs + ky], gzz = G[2 * gs + ky], gxy = G[3 * gs + ky];
                const C2<T> mx = res[ky], my = res[py + ky], mz = res[2 * py + ky];
                const C2<T> hx = cmul(gxx, mx), hx2 = cmul(gxy, my), hy = cmul(gxy, mx), hy2 = cmul(gyy, my);
                const C2<T> hz = cmul(gzz, mz);
                res[ky] = {-(hx.x + hx2.x), -(hx.y + hx2.y)};
                res[py + ky] = {-(hy.x + hy2.x), -(hy.y + hy2.y)};
                res[2 * py + ky] = {-hz.x, -hz.y};
            }
        } else
        for (int ky = threadIdx.x; ky < py; ky += blockDim.x) {
            const bool lo = ky <= ny;
            const int kr = lo ? ky : py - ky;
            const T nxx = N[kr], nyy = N[cs + kr], nzz = N[2 * cs + kr];
            const T nxy = lo ? N[3 * cs + kr] : -N[3 * cs + kr];
            const C2<T> mx = res[ky], my = res[py + ky], mz = res[2 * py + ky];
            res[ky] = {-(nxx * mx.x + nxy * my.x), -(nxx * mx.y + nxy * my.y)};
            res[py + ky] = {-(nxy * mx.x + nyy * my.x), -(nxy * mx.y + nyy * my.y)};
            res[2 * py + ky] = {-(nzz * mz.x), -(nzz * mz.y)};
        }
        __syncthreads();
        C2<T> *other = res == b0 ? b1 : b0;
        res = lds_fft<T, true>(res, other, A.fy, tw, 3, py, ny);
        nstore = ny;
    } else if (MODE == 1) {
        res = lds_fft<T, false>(b0, b1, A.fy, tw, 3, ny, py);
        nstore = py;
    } else {
        res = lds_fft<T, true>(b0, b1, A.fy, tw, 3, py, ny);
        nstore = ny;
    }
    for (int i = threadIdx.x; i < 3 * nstore; i += blockDim.x) {
        const int c = i / nstore, yy = i - c * nstore;
        S[(((long long)c * A.nzs + zz) * nkx + kx) * ly + yy] = res[c * py + yy];
    }
}

// Thin-film y pass with a register four-step FFT of length py = R1*R2:
// one work-group per kx column, TS = max(R1,R2) threads per component.
// Forward: thread t1 holds x[t1 + R1 n2] (only n < ny loaded), R2-point DFT in
// registers, twiddle W_py^{t1 k2}, one padded LDS transpose, R1-point DFT ->
// thread k2 holds X[k2 + R2 k1].  The three components meet in LDS for the
// tensor multiply H = -N M; the inverse runs the same steps backwards and
// keeps the first ny outputs.
template <class T, int R1, int R2>
__global__ void __launch_bounds__(3 * (R1 > R2 ? R1 : R2)) k_yf(const Args<T> *__restrict__ AP) {
    const Args<T> &A = AP[blockIdx.y];
    constexpr int TS = R1 > R2 ? R1 : R2;
    constexpr int N = R1 * R2;
    // the twiddle table comes into LDS with the launch's first loads, moved
    // by the first wave (a global->LDS load moves 256 floats with all 64
    // lanes: the table must be a multiple of that, and the work-group at
    // least one full wave)
    // (measured: FP64 26.6 us against 27.7 on the 512 x 512 film; FP32 is
    // faster with the twiddles from global memory, 10.8 us against 12.5)
    constexpr int TWF = 2 * N * (int)sizeof(T) / 4;
    constexpr bool TWL = sizeof(T) == 8 && TWF % 256 == 0 && 3 * TS >= 64;
    if constexpr (!TWL) {
        if (yz_exit_check(A) && ctl_ro(A)->done) return;
    }
    const int ny = A.ny, ly = A.ly, nkx = A.nkx;
    const int kx = blockIdx.x;
    const int c = threadIdx.x / TS, t = threadIdx.x - c * TS;
    C2<T> *lds = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec) + ((long long)c * A.nzs * nkx + kx) * ly;
    const C2<T> *tw4 = reinterpret_cast<const C2<T> *>(A.tw4y);  // tw4[k2*R1 + t1] = W_N^{t1 k2}
    constexpr int P1 = R1 + 1, P2 = R2 + 1;                    // padded LDS rows
    constexpr int TB = (R2 * P1 > R1 * P2 ? R2 * P1 : R1 * P2);  // transpose block per component
    C2<T> *tb = lds + c * TB;
    C2<T> *twl = lds + (3 * TB + 1) / 2 * 2;                   // 16-byte aligned

    // ---- the loads of the launch (FP64: before the exit check; they only
    // read, and the check's scalar round trip then overlaps them): the
    // tensor values of this thread's frequencies (component c needs N_cc and
    // for x, y N_xy, the sign of ky's half applied at the multiply), the
    // forward transform's inputs and the twiddle table
    const long long cs = (long long)nkx * (ny + 1);
    const T *NT = A.ntens + (long long)kx * (ny + 1);
    const int other = c == 2 ? 2 : (c ^ 1);
    T na[R1], nb[R1];
    if (t < R2) {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            const int k = t + R2 * k1;
            const int kr = k <= ny ? k : N - k;
            na[k1] = ldg(NT + c * cs + kr);
            const T b = ldg(NT + 3 * cs + kr);
            nb[k1] = c == 2 ? T(0) : b;
        }
    }
    // forward step 1 inputs: thread t1 < R1 (py = 2 ny: the upper half of
    // every thread's inputs is the zero padding)
    C2<T> v[R2];
    if (t < R1) {
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) v[n2] = n2 < R2 / 2 ? ldg(S + t + R1 * n2) : C2<T>{T(0), T(0)};
    }
    // R1 = R2 with global twiddles: W_N^{t k} is symmetric in (t, k), so the
    // forward twiddles of thread t (t = t1, k = k2) are also its inverse ones
    // (t = k2, k = t1, conjugated); they are loaded here with the launch's
    // other loads (lanes consecutive) and kept in registers, instead of two
    // dependent table reads after the forward and the inverse DFTs
    constexpr bool TWR = !TWL && R1 == R2;
    C2<T> twr[TWR ? R2 : 1];
    if constexpr (TWR) {
        if (t < R1) {
#pragma unroll
            for (int k = 1; k < R2; ++k) twr[k] = ldg(tw4 + k * R1 + t);
        }
    }
    if constexpr (TWL) {
        if (threadIdx.x < 64)
            dma_floats(reinterpret_cast<const float *>(tw4), reinterpret_cast<float *>(twl), TWF);
    }
    if constexpr (TWL) {
        if (yz_exit_check(A) && ctl_ro(A)->done) {
            // the global->LDS loads land before the work-group's LDS is released
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            return;
        }
    }
    // ---- forward step 1: R2-point DFTs, twiddles, transpose
    if (t < R1) DftReg<T, R2, false>::run(v);
    if constexpr (TWL) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }
    if (t < R1) {
#pragma unroll
        for (int k2 = 1; k2 < R2; ++k2)
            v[k2] = cmul(v[k2], TWL ? twl[k2 * R1 + t] : TWR ? twr[k2] : ldg(tw4 + k2 * R1 + t));
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) tb[k2 * P1 + t] = v[k2];
    }
    __syncthreads();
    // ---- forward step 2: thread k2 < R2
    C2<T> u[R1];
    if (t < R2) {
#pragma unroll
        for (int t1 = 0; t1 < R1; ++t1) u[t1] = tb[t * P1 + t1];
        DftReg<T, R1, false>::run(u);
    }
    __syncthreads();
    // ---- tensor multiply: components exchange through LDS in natural order
    C2<T> *ex = lds;  // 3 x N
    if (t < R2) {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) ex[c * N + t + R2 * k1] = u[k1];
    }
    __syncthreads();
    if (t < R2) {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            const int k = t + R2 * k1;
            const C2<T> ms = ex[c * N + k], mo = ex[other * N + k];
            const T b = k <= ny ? nb[k1] : -nb[k1];
            u[k1] = {-(na[k1] * ms.x + b * mo.x), -(na[k1] * ms.y + b * mo.y)};
        }
    }
    __syncthreads();
    // ---- inverse step A: thread k2 < R2 transforms over k1, twiddle conj
    if (t < R2) {
        DftReg<T, R1, true>::run(u);
#pragma unroll
        for (int t1 = 1; t1 < R1; ++t1) {
            C2<T> w = TWL ? twl[t * R1 + t1] : TWR ? twr[t1] : ldg(tw4 + t * R1 + t1);
            w.y = -w.y;
            u[t1] = cmul(u[t1], w);
        }
#pragma unroll
        for (int t1 = 0; t1 < R1; ++t1) tb[t1 * P2 + t] = u[t1];
    }
    __syncthreads();
    // ---- inverse step B: thread t1 < R1 transforms over k2, keeps n < ny
    if (t < R1) {
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[k2] = tb[t * P2 + k2];
        DftReg<T, R2, true>::run(v);
#pragma unroll
        for (int n2 = 0; n2 < R2 / 2; ++n2) stg(S + t + R1 * n2, v[n2]);
    }
}

// Thin-film y pass at py = 1024 (ny = 512) as three steps of 8-point
// register DFTs, 128 lanes per component (384 threads; one work-group per kx
// column).  Index maps: n = b + 64 m with b = t1 + 8 t1' (t1, t1' < 8,
// m < 16; n < 512 <=> m < 8) and k = k2 + 128 k1 with k2 = k2' + 16 k1'.
// Forward: X[k2 + 128 k1] = sum_t1 W_8^{t1 k1} W_1024^{t1 k2}
//   sum_t1' W_8^{t1' k1'} W_128^{t1' k2'} sum_m x[b + 64 m] W_16^{m k2'}.
// Step A (lane pair b, e): the 16-point DFT over m of 8 nonzero inputs is
// two 8-point DFTs, lane e taking outputs k2' = 2j + e of x[m] W_16^{m e};
// step B (lane t1, k2'): 8-point DFT over t1'; step C (lane k2): 8-point DFT
// over t1.  Lane k2 then holds X at its 8 frequencies, multiplies by -N
// (the other in-plane component through LDS) and runs the steps backwards;
// the inverse's 16-point step keeps outputs m < 8 as A + W_16^{-m} B of the
// lane pair's two 8-point DFTs.  About 1,600 VALU instructions per wave
// against 4,500 for the 32 x 32 four-step form (k_yf; the same total over
// three times the waves): the film's y pass is one work-group per column,
// so its time is a wave's dependent chain.
// Twiddles are powers of per-lane bases from a table (Args::tw8,
// [q][lane]: a wave's loads of one entry are contiguous), loaded with the
// inputs.
constexpr int kYf8Ta = 1040, kYf8Tb = 1056;   // LDS regions per component (complex values)
template <class T>
__host__ __device__ constexpr size_t lds_yf8() {
    return 3ull * (kYf8Ta + kYf8Tb) * sizeof(C2<T>);
}
template <class T>
__global__ void __launch_bounds__(384) k_yf8(const Args<T> *__restrict__ AP) {
    const Args<T> &A = AP[blockIdx.y];
    if (yz_exit_check(A) && ctl_ro(A)->done) return;
    constexpr int NY = 512, N = 1024;
    // LDS layouts, both directions (complex values): step A/B transposes
    // at ta(t1, t1', k2') = 4 t1 + (k2' & 3) + 32 (k2' >> 2) + 130 t1' and
    // step B/C ones at tb(k2, t1) = k2 + 132 t1; both make the 32 lanes of
    // each half-wave hit 32 different bank pairs on either side
    constexpr int TBS = 132;
    auto ta = [](int t1, int t1p, int k2p) { return 4 * t1 + (k2p & 3) + 32 * (k2p >> 2) + 130 * t1p; };
    const int nkx = A.nkx, kx = blockIdx.x;
    const int c = (int)threadIdx.x >> 7, l = (int)threadIdx.x & 127;
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec) + ((long long)c * A.nzs * nkx + kx) * A.ly;
    const C2<T> *W = reinterpret_cast<const C2<T> *>(A.tw8);
    C2<T> *TA = reinterpret_cast<C2<T> *>(fm_smem) + c * kYf8Ta;
    C2<T> *TB = reinterpret_cast<C2<T> *>(fm_smem) + 3 * kYf8Ta + c * kYf8Tb;
    const int other = c == 2 ? 2 : (c ^ 1);
    const C2<T> *TAo = reinterpret_cast<C2<T> *>(fm_smem) + other * kYf8Ta;
    // lane roles: step A/A' (b, e), b = t1 + 8 t1'; step B/B' (t1 = l & 7,
    // k2' = l >> 3); step C/C' k2 = l
    const int e = l & 1, b = l >> 1, at1 = b & 7, at1p = b >> 3;
    const int bt1 = l & 7, bk2p = l >> 3;
    const int k2 = l;

    // ---- the launch's loads: inputs (m = 2i + e), tensor values of the
    // lane's frequencies k = k2 + 128 k1, twiddles of every step
    C2<T> in4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) in4[i] = ldg(S + b + 64 * (2 * i + e));
    const long long cs = (long long)nkx * (NY + 1);
    const T *NT = A.ntens + (long long)kx * (NY + 1);
    T na[8], nb[8];
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) {
        const int k = k2 + 128 * k1;
        const int kr = k <= NY ? k : N - k;
        na[k1] = ldg(NT + c * cs + kr);
        const T bv = ldg(NT + 3 * cs + kr);
        nb[k1] = c == 2 ? T(0) : bv;
    }
    // twiddles as powers of per-lane bases (one table row each):
    // wa[j] = W^{8 t1' e} (W^{16 t1'})^j, wb[j] = W^{t1 k2'} (W^{16 t1})^j,
    // wc[j] = (W^{k2})^j, wd[j] = (W^{8 k2'})^j, W = W_1024
    C2<T> wa[8], wb[8], wc[8], wd[8];
    {
        const C2<T> a0 = ldg(W + 0 * 128 + l), a1 = ldg(W + 1 * 128 + l);
        const C2<T> b0 = ldg(W + 2 * 128 + l), b1 = ldg(W + 3 * 128 + l);
        const C2<T> c1 = ldg(W + 4 * 128 + l), d1 = ldg(W + 5 * 128 + l);
        wa[0] = a0;
        wb[0] = b0;
        wc[0] = {T(1), T(0)};
        wd[0] = {T(1), T(0)};
        wc[1] = c1;
        wd[1] = d1;
#pragma unroll
        for (int j = 1; j < 8; ++j) {
            wa[j] = cmul(wa[j - 1], a1);
            wb[j] = cmul(wb[j - 1], b1);
            if (j > 1) {
                wc[j] = cmul(wc[j - 1], c1);
                wd[j] = cmul(wd[j - 1], d1);
            }
        }
    }

    // ---- forward step A
    C2<T> v[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const C2<T> p = {pair_swap(in4[i].x), pair_swap(in4[i].y)};
        v[2 * i] = e ? p : in4[i];
        v[2 * i + 1] = e ? in4[i] : p;
    }
#pragma unroll
    for (int m = 1; m < 8; ++m) {
        const C2<T> w16 = {T(TwHolder<16>::tw.c[m]), T(-TwHolder<16>::tw.s[m])};
        const C2<T> r = cmul(v[m], w16);
        v[m] = e ? r : v[m];
    }
    DftReg<T, 8, false>::run(v);
#pragma unroll
    for (int j = 0; j < 8; ++j) TA[ta(at1, at1p, 2 * j + e)] = cmul(v[j], wa[j]);
    __syncthreads();
    // ---- forward step B
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = TA[ta(bt1, j, bk2p)];
    DftReg<T, 8, false>::run(v);
#pragma unroll
    for (int j = 0; j < 8; ++j) TB[bk2p + 16 * j + TBS * bt1] = cmul(v[j], wb[j]);
    __syncthreads();
    // ---- forward step C, then X at k2 + 128 k1 to LDS for the exchange
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = TB[k2 + TBS * j];
    DftReg<T, 8, false>::run(v);
#pragma unroll
    for (int j = 0; j < 8; ++j) TA[k2 + 128 * j] = v[j];
    __syncthreads();
    // ---- tensor multiply H = -N M (N_xy's sign flips in ky's upper half)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = k2 + 128 * j;
        const C2<T> mo = TAo[k];
        const T bv = k <= NY ? nb[j] : -nb[j];
        v[j] = {-(na[j] * v[j].x + bv * mo.x), -(na[j] * v[j].y + bv * mo.y)};
    }
    // ---- inverse step C': 8-point over k1 -> t1, conj W_1024^{t1 k2}
    DftReg<T, 8, true>::run(v);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const C2<T> w = {wc[j].x, -wc[j].y};
        TB[k2 + TBS * j] = j ? cmul(v[j], w) : v[j];
    }
    __syncthreads();
    // ---- inverse step B': 8-point over k1' -> t1', conj W_128^{t1' k2'}
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = TB[bk2p + 16 * j + TBS * bt1];
    DftReg<T, 8, true>::run(v);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const C2<T> w = {wd[j].x, -wd[j].y};
        TA[ta(bt1, j, bk2p)] = j ? cmul(v[j], w) : v[j];
    }
    __syncthreads();
    // ---- inverse step A': lane e transforms k2' = 2j + e; the pair adds
    // A + W_16^{-m} B for the outputs m < 8 and stores m = 2i + e
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = TA[ta(at1, at1p, 2 * j + e)];
    DftReg<T, 8, true>::run(v);
#pragma unroll
    for (int m = 1; m < 8; ++m) {
        const C2<T> w16 = {T(TwHolder<16>::tw.c[m]), T(TwHolder<16>::tw.s[m])};
        const C2<T> r = cmul(v[m], w16);
        v[m] = e ? r : v[m];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const C2<T> a0 = v[2 * i], a1 = v[2 * i + 1];
        const C2<T> p0 = {pair_swap(a0.x), pair_swap(a0.y)};
        const C2<T> p1 = {pair_swap(a1.x), pair_swap(a1.y)};
        // even lane: x[2i] = own + partner's; odd lane: x[2i + 1]
        const C2<T> o = e ? a1 : a0, q = e ? p1 : p0;
        stg(S + b + 64 * (2 * i + e), e ? cadd(q, o) : cadd(o, q));
    }
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_z(const Args<T> *__restrict__ AP) {
    const Args<T> &A = AP[blockIdx.y];
    if (ctl_ro(A)->done) return;
    const int pz = A.pz, nz = A.nz, py = A.py, ny = A.ny, ly = A.ly, nkx = A.nkx, TY = A.ty;
    const int ntile = (py + TY - 1) / TY;
    const int kx = blockIdx.x / ntile, ky0 = (blockIdx.x % ntile) * TY;
    C2<T> *b0 = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *b1 = b0 + 3 * TY * pz;
    C2<T> *tw = b1 + 3 * TY * pz;
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec);
    load_twiddles<T>(tw, A.twz, pz);
    for (int i = threadIdx.x; i < 3 * nz * TY; i += blockDim.x) {
        const int t = i % TY, rest = i / TY;
        const int zz = rest % nz, c = rest / nz;
        const int ky = ky0 + t;
        b0[(c * TY + t) * pz + zz] =
            ky < py ? S[(((long long)c * A.nzs + zz) * nkx + kx) * ly + ky] : C2<T>{T(0), T(0)};
    }
    __syncthreads();
    C2<T> *res = lds_fft<T, false>(b0, b1, A.fz, tw, 3 * TY, nz, pz);
    for (int i = threadIdx.x; i < TY * pz; i += blockDim.x) {
        const int t = i / pz, kz = i - t * pz;
        const int ky = ky0 + t;
        if (ky >= py) continue;
        if (A.gtens) {
            // complex spectrum of a periodic kernel: [kx][kz][ky][c]
            const C2<T> *G = reinterpret_cast<const C2<T> *>(A.gtens) + (((size_t)kx * pz + kz) * py + ky) * 6;
            C2<T> &ex = res[(0 * TY + t) * pz + kz];
            C2<T> &ey = res[(1 * TY + t) * pz + kz];
            C2<T> &ez = res[(2 * TY + t) * pz + kz];
            const C2<T> mx = ex, my = ey, mz = ez;
            auto dot3 = [](C2<T> a, C2<T> u, C2<T> b, C2<T> v, C2<T> c, C2<T> w) {
                const C2<T> p = cmul(a, u), q = cmul(b, v), r = cmul(c, w);
                return C2<T>{-(p.x + q.x + r.x), -(p.y + q.y + r.y)};
            };
            ex = dot3(G[0], mx, G[3], my, G[4], mz);
            ey = dot3(G[3], mx, G[1], my, G[5], mz);
            ez = dot3(G[4], mx, G[5], my, G[2], mz);
            continue;
        }
        const bool ylo = ky <= ny, zlo = kz <= nz;
        const int kyr = ylo ? ky : py - ky, kzr = zlo ? kz : pz - kz;
        const T sy = ylo ? T(1) : T(-1), sz = zlo ? T(1) : T(-1);
        const T *N = A.ntens + (((size_t)kx * (nz + 1) + kzr) * (ny + 1) + kyr) * 6;
        const T nxx = N[0], nyy = N[1], nzz = N[2];
        const T nxy = sy * N[3], nxz = sz * N[4], nyz = sy * sz * N[5];
        C2<T> &ex = res[(0 * TY + t) * pz + kz];
        C2<T> &ey = res[(1 * TY + t) * pz + kz];
        C2<T> &ez = res[(2 * TY + t) * pz + kz];
        const C2<T> mx = ex, my = ey, mz = ez;
        ex = {-(nxx * mx.x + nxy * my.x + nxz * mz.x), -(nxx * mx.y + nxy * my.y + nxz * mz.y)};
        ey = {-(nxy * mx.x + nyy * my.x + nyz * mz.x), -(nxy * mx.y + nyy * my.y + nyz * mz.y)};
        ez = {-(nxz * mx.x + nyz * my.x + nzz * mz.x), -(nxz * mx.y + nyz * my.y + nzz * mz.y)};
    }
    __syncthreads();
    C2<T> *other = res == b0 ? b1 : b0;
    res = lds_fft<T, true>(res, other, A.fz, tw, 3 * TY, pz, nz);
    for (int i = threadIdx.x; i < 3 * nz * TY; i += blockDim.x) {
        const int t = i % TY, rest = i / TY;
        const int zz = rest % nz, c = rest / nz;
        const int ky = ky0 + t;
        if (ky < py) S[(((long long)c * A.nzs + zz) * nkx + kx) * ly + ky] = res[(c * TY + t) * pz + zz];
    }
}

// Multilayer demag in the spectral plane (crates/fullmag-engine/src/
// multilayer.rs:166-260): after the x and y transforms of every plane, each
// (kx, ky) column holds the L planes' spectra; plane l receives
// H_l = -sum_q K_{l q} M_q with the pair kernels of the layer heights.  One
// thread per column, the column's M in registers.
template <class T>
__global__ void __launch_bounds__(kThreads) k_lmix(const Args<T> *__restrict__ AP) {
    const Args<T> &A = AP[blockIdx.y];
    if (ctl_ro(A)->done) return;
    const int py = A.py, ny = A.ny, ly = A.ly, nkx = A.nkx, L = A.nz;
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long long)nkx * py) return;
    const int kx = (int)(col / py), ky = (int)(col - (long long)kx * py);
    const bool lo = ky <= ny;
    const int kr = lo ? ky : py - ky;
    const T sy = lo ? T(1) : T(-1);
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec);
    const size_t cs = (size_t)A.nzs * nkx * ly;
    C2<T> mv[3][kMaxLayers];
#pragma unroll
    for (int q = 0; q < kMaxLayers; ++q)
        if (q < L) {
#pragma unroll
            for (int c = 0; c < 3; ++c) mv[c][q] = S[c * cs + ((size_t)q * nkx + kx) * ly + ky];
        }
    const size_t ps = (size_t)nkx * (ny + 1);
    const size_t at = (size_t)kx * (ny + 1) + kr;
#pragma unroll
    for (int l = 0; l < kMaxLayers; ++l) {
        if (l >= L) continue;
        T h[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int q = 0; q < kMaxLayers; ++q) {
            if (q >= L) continue;
            const T *K = A.lmix + (size_t)(l * L + q) * 6 * ps + at;
            const T xx = K[0], yy = K[ps], zz = K[2 * ps], xy = sy * K[3 * ps], xz = K[4 * ps], yz = sy * K[5 * ps];
            const C2<T> mx = mv[0][q], my = mv[1][q], mz = mv[2][q];
            // xz and yz are i times the stored values
            h[0] += xx * mx.x + xy * my.x - xz * mz.y;
            h[1] += xx * mx.y + xy * my.y + xz * mz.x;
            h[2] += xy * mx.x + yy * my.x - yz * mz.y;
            h[3] += xy * mx.y + yy * my.y + yz * mz.x;
            h[4] += zz * mz.x - xz * mx.y - yz * my.y;
            h[5] += zz * mz.y + xz * mx.x + yz * my.x;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) S[c * cs + ((size_t)l * nkx + kx) * ly + ky] = C2<T>{-h[2 * c], -h[2 * c + 1]};
    }
}

// 3-D y passes with a register four-step FFT of length py = R1*R2: each
// group of TS = max(R1,R2) threads owns one (component, z, kx) y-line of the
// spectrum.  Forward: the first ny = py/2 points are read (the rest is the
// zero padding), the full padded spectrum is written.  Inverse: py points
// are read, the first ny written.  The forward pass is launched with the
// factors of the inverse exchanged, so that in both directions the step
// with every thread active is the one that reads HBM.
template <class T, int R1, int R2, bool INV>
__global__ void __launch_bounds__(kThreads) k_yl(const Args<T> *__restrict__ AP) {
    const Args<T> &A = AP[blockIdx.y];
    if (yz_exit_check(A) && ctl_ro(A)->done) return;
    // lines per work-group from the launch (Args::yl_threads, tuned)
    constexpr int TS = R1 > R2 ? R1 : R2;
    const int LPB = (int)blockDim.x / TS;
    constexpr int P1 = R1 + 1, P2 = R2 + 1;
    constexpr int TB = (R2 * P1 > R1 * P2 ? R2 * P1 : R1 * P2);
    const int nlines = 3 * A.nzs * A.nkx;
    const int g = threadIdx.x / TS, t = threadIdx.x - g * TS;
    const int line = blockIdx.x * LPB + g;
    const bool on = line < nlines;
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec) + (size_t)(on ? line : 0) * A.ly;
    const C2<T> *tw4 = reinterpret_cast<const C2<T> *>(INV ? A.tw4y : A.tw4y_f);
    C2<T> *tb = reinterpret_cast<C2<T> *>(fm_smem) + g * TB;
    if constexpr (!INV) {
        if (t < R1) {
            C2<T> v[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) v[n2] = n2 < R2 / 2 ? ldg(S + t + R1 * n2) : C2<T>{T(0), T(0)};
            DftReg<T, R2, false>::run(v);
#pragma unroll
            for (int k2 = 1; k2 < R2; ++k2) v[k2] = cmul(v[k2], ldg(tw4 + tw_fwd_gidx<R1, R2>(t, k2)));
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) tb[k2 * P1 + t] = v[k2];
        }
        __syncthreads();
        if (t < R2) {
            C2<T> u[R1];
#pragma unroll
            for (int t1 = 0; t1 < R1; ++t1) u[t1] = tb[t * P1 + t1];
            DftReg<T, R1, false>::run(u);
            if (on) {
#pragma unroll
                for (int k1 = 0; k1 < R1; ++k1) stg(S + t + R2 * k1, u[k1]);
            }
        }
    } else {
        if (t < R2) {
            C2<T> u[R1];
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) u[k1] = ldg(S + t + R2 * k1);
            DftReg<T, R1, true>::run(u);
#pragma unroll
            for (int t1 = 1; t1 < R1; ++t1) {
                C2<T> w = ldg(tw4 + t * R1 + t1);
                w.y = -w.y;
                u[t1] = cmul(u[t1], w);
            }
#pragma unroll
            for (int t1 = 0; t1 < R1; ++t1) tb[t1 * P2 + t] = u[t1];
        }
        __syncthreads();
        if (t < R1) {
            C2<T> v[R2];
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[k2] = tb[t * P2 + k2];
            DftReg<T, R2, true>::run(v);
            if (on) {
#pragma unroll
                for (int n2 = 0; n2 < R2 / 2; ++n2) stg(S + t + R1 * n2, v[n2]);
            }
        }
    }
}

// 3-D z pass with a register four-step FFT of length pz = R1*R2: a
// work-group takes UB consecutive ky of one kx; the TS = max(R1,R2) threads
// of each ky hold all three components, so the symmetric 6-component real
// tensor multiply H = -N M needs no exchange.  The first nz of the pz
// z-points are read (the rest is the zero padding) and, after the inverse,
// the first nz are written back in place.  The three components pass one
// at a time through one padded LDS transpose [element][ky] (consecutive
// lanes, consecutive ky, hit consecutive addresses), so that a work-group
// needs a third of the LDS and three fit on a CU; the z-point loads of all
// three are issued before the first transform.
template <class T, int R1, int R2, int UB>
__global__ void __launch_bounds__(UB * (R1 > R2 ? R1 : R2)) k_zf(const Args<T> *__restrict__ AP) {
    const Args<T> &A = AP[blockIdx.y];
    // checked once the loads below are issued (they only read)
    const int done = yz_exit_check(A) ? ctl_ro(A)->done : 0;
    constexpr int P1 = R1 + 1, P2 = R2 + 1;
    const int nz = A.nz, py = A.py, ny = A.ny, ly = A.ly, nkx = A.nkx, nzs = A.nzs;
    // ky per work-group from the launch (Args::zf_ub, tuned; at most UB)
    const int ub = (int)blockDim.x / (R1 > R2 ? R1 : R2);
    const int ntile = (py + ub - 1) / ub;
    // (pairing tiles j and ntile-1-j, which read the same tensor records,
    // was measured slower than consecutive tiles)
    const int kx = blockIdx.x / ntile, ky0 = (blockIdx.x - kx * ntile) * ub;
    const int u = threadIdx.x % ub, t = threadIdx.x / ub;
    const int ky = ky0 + u;
    const bool on = ky < py;
    const size_t zs = (size_t)nkx * ly;  // stride between z planes (complex elements)
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec) + (size_t)kx * ly + (on ? ky : py - 1);
    const C2<T> *tw4 = reinterpret_cast<const C2<T> *>(A.tw4z);
    C2<T> *tb = reinterpret_cast<C2<T> *>(fm_smem);
    auto L = [&](int i) -> C2<T> & { return tb[i * ub + u]; };
    // every z-point load of the three components is issued first; thread
    // t < R1 holds z = t + R1 n2 (n2 < R2/2 non-zero)
    C2<T> raw[3][R2 / 2];
    if (t < R1) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int n2 = 0; n2 < R2 / 2; ++n2) raw[c][n2] = ldg(S + (size_t)(c * nzs + t + R1 * n2) * zs);
    }
    if (done) return;
    // forward, one component at a time through one transpose buffer: thread
    // t < R2 ends with kz = t + R2 k1 of every component
    C2<T> x[3][R1];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (t < R1) {
            C2<T> v[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) v[n2] = n2 < R2 / 2 ? raw[c][n2] : C2<T>{T(0), T(0)};
            DftReg<T, R2, false>::run(v);
#pragma unroll
            for (int k2 = 1; k2 < R2; ++k2) v[k2] = cmul(v[k2], ldg(tw4 + tw_fwd_gidx<R1, R2>(t, k2)));
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) L(k2 * P1 + t) = v[k2];
        }
        __syncthreads();
        if (t < R2) {
#pragma unroll
            for (int t1 = 0; t1 < R1; ++t1) x[c][t1] = L(t * P1 + t1);
            DftReg<T, R1, false>::run(x[c]);
        }
        __syncthreads();
    }
    if (t < R2) {
        const bool ylo = ky <= ny;
        const int kyr = ylo ? ky : (on ? py - ky : 0);
        const T sy = ylo ? T(1) : T(-1);
        // tensor [kx][kz][ky][6]: consecutive ky (lanes) read consecutive records
        const T *NT = A.ntens + (size_t)kx * (nz + 1) * (ny + 1) * 6 + (size_t)kyr * 6;
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            const int kz = t + R2 * k1;
            const bool zlo = kz <= nz;
            const int kzr = zlo ? kz : 2 * nz - kz;
            const T sz = zlo ? T(1) : T(-1);
            const gp<const T> N = G(NT) + (size_t)kzr * (ny + 1) * 6;
            const T nxx = N[0], nyy = N[1], nzz = N[2];
            const T nxy = sy * N[3], nxz = sz * N[4], nyz = sy * sz * N[5];
            const C2<T> mx = x[0][k1], my = x[1][k1], mz = x[2][k1];
            x[0][k1] = {-(nxx * mx.x + nxy * my.x + nxz * mz.x), -(nxx * mx.y + nxy * my.y + nxz * mz.y)};
            x[1][k1] = {-(nxy * mx.x + nyy * my.x + nyz * mz.x), -(nxy * mx.y + nyy * my.y + nyz * mz.y)};
            x[2][k1] = {-(nxz * mx.x + nyz * my.x + nzz * mz.x), -(nxz * mx.y + nyz * my.y + nzz * mz.y)};
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            DftReg<T, R1, true>::run(x[c]);
#pragma unroll
            for (int t1 = 1; t1 < R1; ++t1) {
                C2<T> w = ldg(tw4 + t * R1 + t1);
                w.y = -w.y;
                x[c][t1] = cmul(x[c][t1], w);
            }
        }
    }
    // inverse step B per component: thread t < R1 holds z = t + R1 n2, the
    // first nz kept
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (t < R2) {
#pragma unroll
            for (int t1 = 0; t1 < R1; ++t1) L(t1 * P2 + t) = x[c][t1];
        }
        __syncthreads();
        if (t < R1) {
            C2<T> v[R2];
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[k2] = L(t * P2 + k2);
            DftReg<T, R2, true>::run(v);
            if (on) {
#pragma unroll
                for (int n2 = 0; n2 < R2 / 2; ++n2) stg(S + (size_t)(c * nzs + t + R1 * n2) * zs, v[n2]);
            }
        }
        if (c < 2) __syncthreads();
    }
}

// ------------------------------------------------------------ adjoint --
// Reverse sweep of one fixed step of an explicit normalised Runge-Kutta
// method (RK4, Heun; the forward rows are the method table's).  Forward:
// y_1 = m_n, k_j = f(y_j, t_n + c_j dt), y_s = N(v_s) with
// v_s = m_n + dt*scale_s*sum_q a_sq k_q and N(v) = v/|v|, m_{n+1} = y_nstage;
// f(m) = -gb (P m x H + alpha m x (m x H)), H = H_ex + H_d + H_ani + H_ext.
//
// Cotangents, for a cotangent r of y_s and a of k_j:
//   N:  c = (r - y (y.r)) / |v|   ->  cot m_n += c,  cot k_q += dt scale_s a_sq c
//   f:  beta = -gb (P a x m + alpha (a x m) x m)              (= (df/dH)^T a)
//       cot y_j = -gb (P H x a + alpha ((m x H) x a + H x (a x m)))
//               + (2A/mu0) sum_nbr (beta_j/Ms_j - beta_i/Ms_i)/d^2   (exchange Laplacian, symmetric)
//               + Ms_i (-K beta)_i                                  (demag, K symmetric)
//               + (c1 + 3 c2 (m.u)^2)(beta.u) u                     (uniaxial anisotropy)
//   parameters: dJ/dMs_i += -beta_i.(H_ex,i + H_ani,i)/Ms_i + m_i.(-K beta)_i
//               dJ/dH1   += sum_i beta_i f(t) w_i   (time-dependent Zeeman amplitude)
// (-K beta) is the demag convolution of beta with unit Ms: the forward FFT
// passes applied to beta.

template <class T>
__device__ __forceinline__ void cross3(const T a[3], const T b[3], T o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Binds the record of the step being reversed (n = *J.step, slot n mod seg)
// and its start time.
template <class T>
__device__ __forceinline__ AdjArgs<T> adj_bind(const AdjArgs<T> &J0) {
    AdjArgs<T> J = J0;
    const long long n = *J0.step;
    const size_t i = (size_t)(n % J0.seg);
    J.rec_y = J0.rec_y0 + i * (size_t)J0.sv;
    J.rec_h = J0.rec_h0 + i * (size_t)J0.sv;
    J.rec_vn = J0.rec_vn0 + i * (size_t)J0.ss;
    J.t_n = J0.tn[n];
    return J;
}

// beta of stage j at `cell` from the recorded state and the cotangent of k_j
// (FP64 arithmetic for both storage precisions: the reverse sweep sums
// small differences of cotangents over many steps).
template <class T>
__device__ __forceinline__ void adj_beta_cell(const Args<T> &A, const AdjArgs<T> &J, int j, int cell, double b[3]) {
    const int N = (int)A.ncell;
    const T *y = J.rec_y + (size_t)(j - 1) * 3 * N;
    const double *a = J.cot_k + (size_t)j * 3 * N;
    const double m[3] = {(double)y[cell], (double)y[N + cell], (double)y[2 * N + cell]};
    const double av[3] = {a[cell], a[N + cell], a[2 * N + cell]};
    double axm[3], axmxm[3];
    cross3(av, m, axm);
    cross3(axm, m, axmxm);
    const double gb = (double)A.gamma_bar, al = (double)A.alpha;
#pragma unroll
    for (int c = 0; c < 3; ++c) b[c] = -gb * ((A.precession ? axm[c] : 0.0) + al * axmxm[c]);
}

// Cotangent of y_j at `cell` (hd = (-K beta)_i, zero without demag),
// parameter gradients, then the normalisation of stage j and the
// accumulation into cot m_n and cot k_q.  FP64 arithmetic on the stored
// state, field and demag values.
template <class T>
__device__ __forceinline__ void adj_local_cell(const Args<T> &A, const AdjArgs<T> &J, const Method &M, int j,
                                               int cell, int x, int y, int z, const double hd_in[3], T wf,
                                               double amp[3]) {
    const int N = (int)A.ncell;
    if (A.mask && !G(A.mask)[cell]) return;
    const T *yv = J.rec_y + (size_t)(j - 1) * 3 * N;
    const T *Hv = J.rec_h + (size_t)(j - 1) * 3 * N;
    const double *a = J.cot_k + (size_t)j * 3 * N;
    const double m[3] = {(double)yv[cell], (double)yv[N + cell], (double)yv[2 * N + cell]};
    const double H[3] = {(double)Hv[cell], (double)Hv[N + cell], (double)Hv[2 * N + cell]};
    const double av[3] = {a[cell], a[N + cell], a[2 * N + cell]};
    const double bi[3] = {J.beta[cell], J.beta[N + cell], J.beta[2 * N + cell]};
    const double hd[3] = {hd_in[0], hd_in[1], hd_in[2]};
    const double ms_i = A.ms_cell ? (double)G(A.ms_cell)[cell] : (double)A.ms;
    const double gb = (double)A.gamma_bar, al = (double)A.alpha;
    double Hxa[3], mxH[3], mxHxa[3], axm[3], Hxaxm[3];
    cross3(H, av, Hxa);
    cross3(m, H, mxH);
    cross3(mxH, av, mxHxa);
    cross3(av, m, axm);
    cross3(H, axm, Hxaxm);
    double r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = -gb * ((A.precession ? Hxa[c] : 0.0) + al * (mxHxa[c] + Hxaxm[c]));
    // exchange: adjoint stencil on beta/Ms, and H_ex at y_j for the Ms gradient
    const int sy = A.nx, sz = A.nx * A.ny;
    const int nb[6] = {x > 0 ? cell - 1 : cell, x < A.nx - 1 ? cell + 1 : cell,
                       y > 0 ? cell - sy : cell, y < A.ny - 1 ? cell + sy : cell,
                       z > 0 ? cell - sz : cell, z < A.nz - 1 ? cell + sz : cell};
    const double wq[6] = {(double)A.idx2, (double)A.idx2, (double)A.idy2, (double)A.idy2, (double)A.idz2,
                          (double)A.idz2};
    const double c_ex = (double)A.two_a_mu0;
    double hex[3] = {0.0, 0.0, 0.0};
    const double inv_ms_i = 1.0 / ms_i;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int o = nb[q];
        const bool use = o != cell && (!A.mask || G(A.mask)[o]);
        if (!use) continue;
        const double inv_ms_o = A.ms_cell ? 1.0 / (double)G(A.ms_cell)[o] : inv_ms_i;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r[c] += c_ex * wq[q] * (J.beta[c * N + o] * inv_ms_o - bi[c] * inv_ms_i);
            hex[c] += wq[q] * ((double)yv[c * N + o] - m[c]);
        }
    }
    const double pre = A.ms_cell ? c_ex * inv_ms_i : (double)A.ex_pre;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        hex[c] *= pre;
        r[c] += ms_i * hd[c];
    }
    double hani[3] = {0.0, 0.0, 0.0};
    if (A.has_uni) {
        double u[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = A.axis_cell ? (double)G(A.axis_cell)[c * N + cell] : (double)A.axis[c];
        double c1 = (double)A.uni_c1, c2 = (double)A.uni_c2;
        if (A.ku1_cell || A.ku2_cell || A.ms_cell) {
            const double k1 = A.ku1_cell ? (double)G(A.ku1_cell)[cell] : (double)A.ku1_u;
            const double k2 = A.ku2_cell ? (double)G(A.ku2_cell)[cell] : (double)A.ku2_u;
            c1 = (double)A.two_mu0inv * k1 * inv_ms_i;
            c2 = 2.0 * (double)A.two_mu0inv * k2 * inv_ms_i;
        }
        const double mu = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
        const double bu = bi[0] * u[0] + bi[1] * u[1] + bi[2] * u[2];
        const double dcoef = (c1 + 3.0 * c2 * mu * mu) * bu;
        const double coef = c1 * mu + c2 * mu * mu * mu;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r[c] += dcoef * u[c];
            hani[c] = coef * u[c];
        }
    }
    // parameter gradients
    {
        double g = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) g += -bi[c] * (hex[c] + hani[c]) * inv_ms_i + m[c] * hd[c];
        J.grad_ms[cell] += g;
        if (A.wave_kind && (!A.wave_mask || G(A.wave_mask)[cell])) {
#pragma unroll
            for (int c = 0; c < 3; ++c) amp[c] += bi[c] * (double)wf;
        }
    }
    if (j >= 2) {
        const double vn = (double)J.rec_vn[(size_t)(j - 2) * N + cell];
        const double yr = m[0] * r[0] + m[1] * r[1] + m[2] * r[2];
        double cv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) cv[c] = vn > 0.0 ? (r[c] - m[c] * yr) / vn : 0.0;
        const StageRow &R = M.row[j];
        const double f = J.dt * R.scale;
#pragma unroll
        for (int c = 0; c < 3; ++c) J.cot_m[c * N + cell] += cv[c];
        for (int q = 0; q < R.nk; ++q) {
            double *ck = J.cot_k + (size_t)R.kidx[q] * 3 * N;
            const double w = f * R.a[q];
#pragma unroll
            for (int c = 0; c < 3; ++c) ck[c * N + cell] += w * cv[c];
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) J.cot_m[c * N + cell] += r[c];
    }
}

// Deterministic block sum of 3 doubles added to out[0..3) by thread 0.
__device__ void block_add3(double v[3], double *out) {
    __shared__ double red3[3][16];  // up to 1024 threads
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // the three components' exchanges per distance together (one exchange
    // latency per distance; per-component order unchanged)
    double x[3] = {v[0], v[1], v[2]};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double y[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) y[i] = __shfl_xor(x[i], off, 64);
#pragma unroll
        for (int i = 0; i < 3; ++i) x[i] += y[i];
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) red3[i][wid] = x[i];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double x = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) x += red3[threadIdx.x][w];
        out[threadIdx.x] += x;
    }
}

// Start of the reverse sweep of step n: lambda = cot m_{n+1} (+ objective
// derivative at m_{n+1}), normalisation of the final stage, k cotangents
// of the final row; cot m_n restarts from there.
template <class T>
__global__ void __launch_bounds__(kThreads) k_adj_init(const Args<T> *__restrict__ AP, AdjArgs<T> J0, int nstage) {
    const Args<T> &A = AP[blockIdx.y];
    const int N = (int)A.ncell;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= N) return;
    const AdjArgs<T> J = adj_bind(J0);
    const bool with_obj = *J0.step + 1 >= J0.first;
    const Method &M = c_methods[A.method];
    // the step's end state m_{n+1} starts the next slot of the record ring
    const T *y = J.rec_y0 + (size_t)((*J0.step + 1) % J0.seg) * (size_t)J0.sv;
    const double m[3] = {(double)y[cell], (double)y[N + cell], (double)y[2 * N + cell]};
    double lam[3] = {J.cot_m[cell], J.cot_m[N + cell], J.cot_m[2 * N + cell]};
    if (with_obj) lam[J.obj_comp] += 2.0 * J.obj_scale * (double)J.obj_w[cell] * m[J.obj_comp];
    const double vn = (double)J.rec_vn[(size_t)(nstage - 2) * N + cell];
    const double yl = m[0] * lam[0] + m[1] * lam[1] + m[2] * lam[2];
    double cv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) cv[c] = vn > 0.0 ? (lam[c] - m[c] * yl) / vn : 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) J.cot_m[c * N + cell] = cv[c];
    for (int j = 1; j < nstage; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) J.cot_k[((size_t)j * 3 + c) * N + cell] = 0.0;
    const StageRow &R = M.row[nstage];
    const double f = J.dt * R.scale;
    for (int q = 0; q < R.nk; ++q) {
        double *ck = J.cot_k + (size_t)R.kidx[q] * 3 * N;
        const double w = f * R.a[q];
#pragma unroll
        for (int c = 0; c < 3; ++c) ck[c * N + cell] += w * cv[c];
    }
}

// Objective partial sums of the current state: obj_part[block] += sum w_i m_c,i^2.
template <class T>
__global__ void __launch_bounds__(kThreads) k_adj_objective(const Args<T> *__restrict__ AP, AdjArgs<T> J) {
    const Args<T> &A = AP[blockIdx.y];
    if ((long long)A.ctl->step < J.obj_from) return;
    const int N = (int)A.ncell;
    const T *m = A.m_all + (size_t)A.ctl->parity * 3 * N + (size_t)J.obj_comp * N;
    double s = 0.0;
    for (int cell = blockIdx.x * blockDim.x + threadIdx.x; cell < N; cell += gridDim.x * blockDim.x)
        s += (double)J.obj_w[cell] * (double)m[cell] * (double)m[cell];
    double v[3] = {s, 0.0, 0.0};
    block_add3(v, J.obj_part + (size_t)blockIdx.x * 3);
}

// Sets (add = 0) or steps (add != 0) the device step index of the sweep.
static __global__ void k_adj_step(long long *step, long long value, int add) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step = add ? *step + value : value;
}

// Without demag: beta of stage j, one thread per cell.
template <class T>
__global__ void __launch_bounds__(kThreads) k_adj_beta(const Args<T> *__restrict__ AP, AdjArgs<T> J0, int j) {
    const Args<T> &A = AP[blockIdx.y];
    const AdjArgs<T> J = adj_bind(J0);
    const int N = (int)A.ncell;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= N) return;
    double b[3];
    adj_beta_cell(A, J, j, cell, b);
#pragma unroll
    for (int c = 0; c < 3; ++c) J.beta[c * N + cell] = b[c];
}

// Without demag: cotangent of y_j, one thread per cell.
template <class T>
__global__ void __launch_bounds__(kThreads) k_adj_local(const Args<T> *__restrict__ AP, AdjArgs<T> J0, int j) {
    const Args<T> &A = AP[blockIdx.y];
    const AdjArgs<T> J = adj_bind(J0);
    const int N = (int)A.ncell;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    const Method &M = c_methods[A.method];
    double amp[3] = {0.0, 0.0, 0.0};
    if (cell < N) {
        const int x = cell % A.nx, y = (cell / A.nx) % A.ny, z = cell / (A.nx * A.ny);
        const double hd[3] = {0.0, 0.0, 0.0};
        const T wf = A.wave_kind ? wave_factor(A, J.t_n + M.node[j] * J.dt) : T(0);
        adj_local_cell(A, J, M, j, cell, x, y, z, hd, wf, amp);
    }
    block_add3(amp, J.grad_amp + (size_t)blockIdx.x * 3);
}

// With demag: beta of stage j for the work-group's x-rows, stored and
// packed (unit Ms) two real rows per complex sequence, forward x FFT.  TF is
// the precision of the field cotangent's convolution: FP points to the
// argument record of a demag context of that precision on the same grid
// (the FP64 context of an FP32 system, or AP itself when TF == T); the rows
// of a work-group and the transforms follow it.
template <class T, class TF, int R1, int R2>
__global__ void __launch_bounds__(kThreads) k_adj_fx(const Args<T> *__restrict__ AP, const Args<TF> *__restrict__ FP,
                                                     AdjArgs<T> J0, int j) {
    const Args<T> &A = AP[blockIdx.y];
    const Args<TF> &F = FP[blockIdx.y];
    constexpr bool FAST = R1 > 0;
    const AdjArgs<T> J = adj_bind(J0);
    C2<TF> *b0 = reinterpret_cast<C2<TF> *>(fm_smem);
    C2<TF> *b1 = b0 + F.nseq_x * F.xsb;
    C2<TF> *tw = b1 + F.nseq_x * F.xsb;
    __shared__ XRows xr;
    x_rows(F, (int)xcd_block(blockIdx.x, gridDim.x) * F.rows_per_wg, xr);
    if constexpr (R1 == 0) load_twiddles<TF>(tw, F.twx, F.px);
    __syncthreads();
    const int px = FAST ? R1 * R2 : F.px, nx = px >> 1;
    const int rows = F.rows_per_wg, ncl = rows * nx, nl = 3 * rows, N = (int)A.ncell;
    const int sb = FAST ? F.xsb : px;
    const float rnx = 1.0f / (float)nx;
    for (int i = threadIdx.x; i < ncl; i += blockDim.x) {
        const int rr = small_div<FAST>(i, nx, rnx), x = i - rr * nx;
        const int c0 = xr.cell0[rr];
        double b[3] = {0.0, 0.0, 0.0};
        if (c0 >= 0) {
            adj_beta_cell(A, J, j, c0 + x, b);
#pragma unroll
            for (int c = 0; c < 3; ++c) J.beta[c * N + c0 + x] = b[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int l = 3 * rr + c;
            TF *e = reinterpret_cast<TF *>(&b0[(l >> 1) * sb + x]);
            e[l & 1] = (TF)b[c];
        }
    }
    if (nl & 1)
        for (int x = threadIdx.x; x < nx; x += blockDim.x) b0[(nl >> 1) * sb + x].y = TF(0);
    x_forward_store<TF, R1, R2>(F, xr, b0, b1, tw);
}

// With demag: inverse x FFT of the convolved beta spectrum (1/P), then the
// cotangent of y_j for the work-group's cells (FP as in k_adj_fx).
template <class T, class TF, int R1, int R2>
__global__ void __launch_bounds__(kThreads) k_adj_bx(const Args<T> *__restrict__ AP, const Args<TF> *__restrict__ FP,
                                                     AdjArgs<T> J0, int j) {
    const Args<T> &A = AP[blockIdx.y];
    const Args<TF> &F = FP[blockIdx.y];
    constexpr bool FAST = R1 > 0;
    const AdjArgs<T> J = adj_bind(J0);
    const Method &M = c_methods[A.method];
    C2<TF> *b0 = reinterpret_cast<C2<TF> *>(fm_smem);
    C2<TF> *b1 = b0 + F.nseq_x * F.xsb;
    C2<TF> *tw = b1 + F.nseq_x * F.xsb;
    __shared__ XRows xr;
    x_rows(F, (int)xcd_block(blockIdx.x, gridDim.x) * F.rows_per_wg, xr);
    if constexpr (R1 == 0) load_twiddles<TF>(tw, F.twx, F.px);
    __syncthreads();
    const C2<TF> *res = x_inverse_load<TF, R1, R2>(F, xr, b0, b1, tw);
    const int px = FAST ? R1 * R2 : F.px, nx = px >> 1;
    const int ncl = F.rows_per_wg * nx, sb = FAST ? F.xsb : px;
    const float rnx = 1.0f / (float)nx;
    const double scale = F.inv_p;
    const T wf = A.wave_kind ? wave_factor(A, J.t_n + M.node[j] * J.dt) : T(0);
    double amp[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < ncl; i += blockDim.x) {
        const int rr = small_div<FAST>(i, nx, rnx), x = i - rr * nx;
        const int c0 = xr.cell0[rr];
        if (c0 < 0) continue;
        double hd[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int l = 3 * rr + c;
            const C2<TF> zv = res[(l >> 1) * sb + x];
            hd[c] = (double)(((l & 1) ? zv.y : zv.x) * TF(scale));
        }
        adj_local_cell(A, J, M, j, c0 + x, x, xr.y[rr], xr.z[rr], hd, wf, amp);
    }
    block_add3(amp, J.grad_amp + (size_t)blockIdx.x * 3);
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_stage(const Args<T> *__restrict__ AP, int s) {
    const Args<T> &A = AP[blockIdx.y];
    const CtlR ctl = ctl_ro(A);
    if (ctl->done) return;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= A.ncell) return;
    const Method &M = method_for(A);
    T v[3];
    stage_state(A, M, s, ctl->parity, ctl->dt, cell, v);
}

template <class T, int FT>
__global__ void __launch_bounds__(kThreads) k_local(const Args<T> *__restrict__ AP, int s, int stats) {
    const Args<T> &A = AP[blockIdx.y];
    const CtlR ctl = ctl_ro(A);
    // the control fields in one round trip, then the early exit
    const int done = ctl->done, p = ctl->parity;
    const double dt = ctl->dt;
    if (done) return;
    const Method &M = method_for(A);
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool err = M.adaptive && s == M.nstage;
    Acc acc;
    acc.init();
    if (cell < A.ncell) {
        const int x = (int)(cell % A.nx);
        const int y = (int)((cell / A.nx) % A.ny);
        const int z = (int)(cell / ((long long)A.nx * A.ny));
        const T hd[3] = {T(0), T(0), T(0)};
        const TimeF<T> tf = time_factors(A, stage_time(A, M, s, ctl, dt));
        CellIn<T> g;
        gather_cell<T, FT>(A, M, p, (int)cell, x, y, z, stage_input(A, M, s, p), err, g);
        compute_cell<T, FT>(A, M, (int)cell, dt, g, stage_rhs(A, M, s, p), hd, tf, stats != 0, err, acc,
                             rec_h_ptr(A, M, s),
                             nullptr, stage_push(A, M, s));
    }
    if (stats || err) block_reduce_global(acc, A.partials + (long long)blockIdx.x * kStatSlots);
}

// Reduction of all per-block partials and the step controller, executed by
// one work-group.  observe != 0: only store the statistics.
template <class T>
__device__ void control_block(const Args<T> &A, int nparts, int observe) {
    // the control block, the partials and the recording counter are device
    // allocations: global accesses, which do not share the LDS wait counter
    // (a generic store there made every later LDS read wait for it)
    const gp<Ctl> c = G(A.ctl);
    // every control field the decision reads, and the method's, loaded with
    // the partials (one round trip instead of a chain after the reduction)
    const Method &M = method_for(A);
    const double dt = c->dt, t0 = c->t, abm_dt = c->abm_dt, t_end = c->t_end, dt_carry = c->dt_carry;
    const unsigned long long attempts = c->attempts, rejected = c->rejected, step = c->step, step_end = c->step_end;
    const int parity = c->parity, abm_head = c->abm_head, abm_count = c->abm_count, loop = c->loop;
    const int done = c->done;
    const int adaptive = M.adaptive, abm = M.abm;
    const double ctrl_exp = M.ctrl_exp;
    Acc acc;
    acc.init();
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        // written by the previous launch: the kernel boundary made them visible
        const gp<const double> q = G((const double *)A.partials) + (long long)i * kStatSlots;
#pragma unroll
        for (int k = 0; k < kStatSlots; ++k) {
            const double v = q[k];
            acc.v[k] = slot_is_max(k) ? fmax(acc.v[k], v) : acc.v[k] + v;
        }
    }
    // the device loop has finished: nothing to decide (checked here, so
    // that its round trip overlaps the loads above; uniform per work-group)
    if (done) return;
    __shared__ double tot[kStatSlots];
    block_reduce_lds(acc, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (observe) {
        for (int k = 0; k < kStatSlots; ++k) c->stats[k] = tot[k];
        return;
    }
    c->attempts = attempts + 1;
    bool accept = true;
    double dt_next = 0.0;
    if (adaptive) {
        const double err = tot[ST_ERR];
        c->err = err;
        const double thr = A.rtol > 0.0 ? 1.0 : A.atol;
        accept = err <= thr || dt <= A.dt_min;
        const double ratio =
            fmax(fmin(A.headroom * pow(thr / (accept ? fmax(err, 1e-30) : err), ctrl_exp), A.growth), A.shrink);
        const double dtn = fmin(fmax(dt * ratio, A.dt_min), A.dt_max);
        if (accept) c->dt_next = dt_next = dtn;
        else c->dt = dtn;
    } else {
        c->dt_next = 0.0;
    }
    c->k1_redo = accept ? 0 : 1;
    if (!accept) {
        c->rejected = rejected + 1;
        c->accepted = 0;
        return;
    }
    for (int k = 0; k < kStatSlots; ++k) c->stats[k] = tot[k];
    if (A.rec_step) *G(A.rec_step) += 1;
    const double t = t0 + dt;
    const unsigned long long stepn = step + 1;
    c->t = t;
    c->step = stepn;
    c->parity = parity ^ 1;
    c->dt_used = dt;
    c->accepted = 1;
    if (abm) {
        // AbmHistory::push (fdm_state.rs:247-257): the kernels stored the
        // pushed value in the slot after the head; a dt change above 10%
        // restarts the history first
        const int cnt = (abm_dt > 0.0 && fabs(dt - abm_dt) / abm_dt > 0.1) ? 0 : abm_count;
        c->abm_head = (abm_head + 1) % 3;
        c->abm_count = cnt < 3 ? cnt + 1 : 3;
        c->abm_dt = dt;
    }
    if (loop) {
        double carry = dt_carry;
        if (adaptive) {
            carry = dt_next;
            c->dt_carry = carry;
        }
        if (t >= t_end || stepn >= step_end) {
            c->done = 1;
        } else {
            double d = fmin(carry, t_end - t);
            if (adaptive) d = fmin(fmax(d, A.dt_min), A.dt_max);
            c->dt = d;
        }
    }
}

// Reduction of the partials and the step controller (observe != 0: only the
// statistics), one work-group per system, launched after the
// partial-producing kernel: the kernel boundary makes every work-group's
// partials visible.  (Fusing it into that kernel's last work-group needs a
// device-scope release in every work-group, which writes back the XCD's L2
// each time; measured 11% of a 256x256x64 step.)
template <class T>
__global__ void __launch_bounds__(kThreads) k_control(const Args<T> *__restrict__ AP, int nparts, int observe) {
    control_block(AP[blockIdx.y], nparts, observe);
}

// ------------------------------------------------------------ host side --
static int g_methods_ready = -1;

static void build_methods(Method *out) {
    for (int i = 0; i < M_COUNT; ++i) out[i] = Method{};
    auto row = [](Method &m, int s, std::initializer_list<int> j, std::initializer_list<double> a, double scale) {
        StageRow &r = m.row[s];
        r.nk = 0;
        auto ja = j.begin();
        auto aa = a.begin();
        for (; ja != j.end(); ++ja, ++aa) {
            r.kidx[r.nk] = *ja;
            r.a[r.nk] = *aa;
            ++r.nk;
        }
        r.scale = scale;
    };
    // Heun (fdm_integrators.rs:84-150)
    Method &h = out[M_HEUN];
    h.nstage = 3;
    h.adaptive = 0;
    h.node[2] = 1.0;
    h.node[3] = 1.0;
    row(h, 2, {1}, {1.0}, 1.0);
    row(h, 3, {1, 2}, {1.0, 1.0}, 0.5);
    // Dormand-Prince 5(4) (fdm_integrators.rs:441-645)
    Method &d = out[M_DP45];
    d.nstage = 7;
    d.adaptive = 1;
    d.ctrl_exp = 0.2;
    const double dnode[8] = {0.0, 0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
    for (int i = 0; i < 8; ++i) d.node[i] = dnode[i];
    row(d, 2, {1}, {1.0}, 1.0 / 5.0);
    row(d, 3, {1, 2}, {3.0 / 40.0, 9.0 / 40.0}, 1.0);
    row(d, 4, {1, 2, 3}, {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0}, 1.0);
    row(d, 5, {1, 2, 3, 4}, {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0}, 1.0);
    row(d, 6, {1, 2, 3, 4, 5}, {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
        1.0);
    row(d, 7, {1, 3, 4, 5, 6}, {35.0 / 384.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}, 1.0);
    const int ej[6] = {1, 3, 4, 5, 6, 7};
    const double ew[6] = {71.0 / 57600.0, -71.0 / 16695.0, 71.0 / 1920.0, -17253.0 / 339200.0, 22.0 / 525.0,
                          -1.0 / 40.0};
    d.nerr = 6;
    for (int i = 0; i < 6; ++i) {
        d.eidx[i] = ej[i];
        d.e[i] = ew[i];
    }
    // classic RK4 with normalised stages (fdm_integrators.rs:155-264)
    Method &r = out[M_RK4];
    r.nstage = 5;
    r.adaptive = 0;
    r.node[2] = 0.5;
    r.node[3] = 0.5;
    r.node[4] = 1.0;
    r.node[5] = 1.0;
    row(r, 2, {1}, {1.0}, 0.5);
    row(r, 3, {2}, {1.0}, 0.5);
    row(r, 4, {3}, {1.0}, 1.0);
    row(r, 5, {1, 2, 3, 4}, {1.0, 2.0, 2.0, 1.0}, 1.0 / 6.0);
    // Bogacki-Shampine 3(2) with normalised stages (fdm_integrators.rs:290-410);
    // the k of the new state is the next step's k1 (the reference evaluates
    // it again at the start of the next step, with the same result)
    Method &b = out[M_RK23];
    b.nstage = 4;
    b.adaptive = 1;
    b.ctrl_exp = 1.0 / 3.0;
    b.node[2] = 0.5;
    b.node[3] = 0.75;
    b.node[4] = 1.0;
    row(b, 2, {1}, {1.0}, 0.5);
    row(b, 3, {2}, {1.0}, 0.75);
    row(b, 4, {1, 2, 3}, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}, 1.0);
    const int bj[4] = {1, 2, 3, 4};
    const double bw[4] = {-5.0 / 72.0, 1.0 / 12.0, 1.0 / 9.0, -1.0 / 8.0};
    b.nerr = 4;
    for (int i = 0; i < 4; ++i) {
        b.eidx[i] = bj[i];
        b.e[i] = bw[i];
    }
    // ABM3 (fdm_integrators.rs:751-820): predictor y_2 = m0 + dt (23 f_n -
    // 16 f_{n-1} + 5 f_{n-2})/12, f* = RHS(y_2), corrector m0 + dt (5 f* +
    // 8 f_n - f_{n-1})/12; both evaluated at t + dt.  The RHS of the new
    // state is evaluated for the step's observables and a later start-up.
    Method &ab = out[M_ABM3];
    ab.nstage = 3;
    ab.abm = 1;
    ab.node[2] = 1.0;
    ab.node[3] = 1.0;
    row(ab, 2, {kHist, kHist + 1, kHist + 2}, {23.0 / 12.0, -16.0 / 12.0, 5.0 / 12.0}, 1.0);
    row(ab, 3, {2, kHist, kHist + 1}, {5.0 / 12.0, 8.0 / 12.0, -1.0 / 12.0}, 1.0);
    // its start-up (fdm_integrators.rs:693-749): the Heun rows, padded with
    // zero-weight terms to the same k-vector counts so that one captured
    // launch sequence serves both tables
    Method &as = out[M_ABM3S];
    as.nstage = 3;
    as.abm = 2;
    as.node[2] = 1.0;
    as.node[3] = 1.0;
    row(as, 2, {1, 1, 1}, {1.0, 0.0, 0.0}, 1.0);
    row(as, 3, {1, 2, 1}, {1.0, 1.0, 0.0}, 0.5);
    for (int i = 0; i < M_COUNT; ++i) {
        Method &m = out[i];
        const StageRow &fr = m.row[m.nstage];
        bool f = m.adaptive && m.nerr >= 1 && m.eidx[m.nerr - 1] == m.nstage && fr.nk == m.nerr - 1;
        for (int q = 0; f && q < fr.nk; ++q) f = fr.kidx[q] == m.eidx[q];
        m.efuse = f ? 1 : 0;
    }
}

template <class T>
hipError_t upload_methods() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (g_methods_ready == dev) return hipSuccess;
    Method tab[M_COUNT];
    build_methods(tab);
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_methods), tab, sizeof(tab));
    if (e == hipSuccess) g_methods_ready = dev;
    return e;
}

#ifdef FM_PRIMARY
int method_nstage(int method) {
    Method tab[M_COUNT];
    build_methods(tab);
    return tab[method].nstage;
}

int method_efuse(int method) {
    Method tab[M_COUNT];
    build_methods(tab);
    return tab[method].efuse;
}

int method_stage_nk(int method, int s) {
    Method tab[M_COUNT];
    build_methods(tab);
    return s >= 2 && s <= tab[method].nstage ? tab[method].row[s].nk : 0;
}
#endif

// Four-step factorisation py = R1*R2 of the thin-film fast y pass.
static bool yfast_shape(int py, int &r1, int &r2) {
    switch (py) {
        case 64: r1 = 8; r2 = 8; return true;
        case 128: r1 = 8; r2 = 16; return true;
        case 256: r1 = 16; r2 = 16; return true;
        case 512: r1 = 16; r2 = 32; return true;
        case 1024: r1 = 32; r2 = 32; return true;
        case 2048: r1 = 64; r2 = 32; return true;
        default: return false;
    }
}

// Launch grid of a (possibly batched) launch: x = work-groups of one
// system, y = systems sharing the launch (each reads its arguments at
// AP[blockIdx.y]).
template <class T>
static dim3 bgrid(const Args<T> &A, dim3 g) {
    return dim3(g.x, (unsigned)A.nsys, 1);
}

// Four-step factorisation pz = R1*R2 of the 3-D fast z pass.
static bool zfast_shape(int pz, int &r1, int &r2) {
    switch (pz) {
        case 4: r1 = 2; r2 = 2; return true;
        case 8: r1 = 4; r2 = 2; return true;
        case 16: r1 = 4; r2 = 4; return true;
        case 32: r1 = 8; r2 = 4; return true;
        case 64: r1 = 8; r2 = 8; return true;
        case 128: r1 = 16; r2 = 8; return true;
        case 256: r1 = 16; r2 = 16; return true;
        default: return false;
    }
}
// ky lines per work-group of the fast z pass: 256 threads in all
template <int R1, int R2>
constexpr int zf_units() {
    return kThreads / (R1 > R2 ? R1 : R2);
}

// Tile sizes of the 3-D y-line and z passes: the tuned value when it is a
// valid tile of the shape (a multiple of the line's threads, at most the
// compiled maximum), else the default (kThreads threads, zf_units ky).
template <int R1, int R2>
int yl_threads_of(int req) {
    constexpr int TS = R1 > R2 ? R1 : R2;
    return (req >= TS && req <= kThreads && req % TS == 0 && req % 64 == 0) ? req : kThreads;
}
template <int R1, int R2>
int zf_ub_of(int req) {
    return (req >= 1 && req <= zf_units<R1, R2>()) ? req : zf_units<R1, R2>();
}

template <class T, int R1, int R2>
static size_t lds_yl(int threads = kThreads) {
    constexpr int TS = R1 > R2 ? R1 : R2;
    constexpr int P1 = R1 + 1, P2 = R2 + 1;
    constexpr int TB = (R2 * P1 > R1 * P2 ? R2 * P1 : R1 * P2);
    return (size_t)(threads / TS) * TB * sizeof(C2<T>);
}
template <class T, int R1, int R2>
static size_t lds_zf(int ub = zf_units<R1, R2>()) {
    constexpr int P1 = R1 + 1, P2 = R2 + 1;
    constexpr int TB = (R2 * P1 > R1 * P2 ? R2 * P1 : R1 * P2);
    return 1ull * TB * ub * sizeof(C2<T>);
}

template <class T>
static size_t lds_x(const Args<T> &A) {
    int r1 = 0, r2 = 0;
    if (A.xfast && yfast_shape(A.xfast, r1, r2)) {
        // sequences + one round of padded transpose blocks (no twiddle table)
        const int ts = r1 > r2 ? r1 : r2;
        const long long tb = (long long)(r2 * (r1 + 1) > r1 * (r2 + 1) ? r2 * (r1 + 1) : r1 * (r2 + 1));
        const long long rounds = A.nseq_x < kThreads / ts ? A.nseq_x : kThreads / ts;
        return ((long long)A.nseq_x * A.xsb + rounds * tb) * sizeof(C2<T>);
    }
    return (2ull * A.nseq_x + 1) * A.px * sizeof(C2<T>);
}
template <class T>
static size_t lds_y(const Args<T> &A) { return (2ull * 3 + 1) * A.py * sizeof(C2<T>); }
template <class T>
static size_t lds_z(const Args<T> &A) { return (2ull * 3 * A.ty + 1) * A.pz * sizeof(C2<T>); }

template <class T>
static hipError_t set_lds(const void *fn, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Chooses work-group tiling for the given LDS limit and sets the kernel
// attributes.  Returns an error string (empty on success).
template <class T, int R1, int R2>
static size_t lds_yf() {
    constexpr int P1 = R1 + 1, P2 = R2 + 1;
    constexpr int TB = (R2 * P1 > R1 * P2 ? R2 * P1 : R1 * P2);
    // transpose blocks (the exchange of the components reuses them), then
    // the twiddle table, 16-byte aligned
    return ((3ull * TB + 1) / 2 * 2 + (size_t)R1 * R2) * sizeof(C2<T>);
}

template <class T, int FT, int R1, int R2>
static hipError_t set_lds_xx(size_t bytes) {
    hipError_t e = set_lds<T>((const void *)k_xx<T, FT, R1, R2, 1>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xx<T, FT, R1, R2, 2>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xx<T, FT, R1, R2, 3>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xx<T, FT, R1, R2, 4>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xx<T, FT, R1, R2, 5>, bytes);
    return e;
}

template <class T, int R1, int R2>
static hipError_t set_lds_x_shape(size_t bytes) {
    hipError_t e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 0>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 1>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 2>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 3>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 4>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 5>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_inv_x<T, 0, R1, R2>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_inv_x<T, 1, R1, R2>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_inv_x<T, 2, R1, R2>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_inv_x<T, 3, R1, R2>, bytes);
    if (e == hipSuccess) e = set_lds_xx<T, 0, R1, R2>(bytes);
    if (e == hipSuccess) e = set_lds_xx<T, 1, R1, R2>(bytes);
    if (e == hipSuccess) e = set_lds_xx<T, 2, R1, R2>(bytes);
    if (e == hipSuccess) e = set_lds_xx<T, 3, R1, R2>(bytes);
    return e;
}
template <class T>
static hipError_t set_lds_x(int xfast, size_t bytes) {
    switch (xfast) {
        case 64: return set_lds_x_shape<T, 8, 8>(bytes);
        case 128: return set_lds_x_shape<T, 8, 16>(bytes);
        case 256: return set_lds_x_shape<T, 16, 16>(bytes);
        case 512: return set_lds_x_shape<T, 16, 32>(bytes);
        case 1024: return set_lds_x_shape<T, 32, 32>(bytes);
        case 2048: return set_lds_x_shape<T, 64, 32>(bytes);
        default: return set_lds_x_shape<T, 0, 0>(bytes);
    }
}

template <class T, int R1, int R2>
static hipError_t set_lds_yl_shape(size_t limit) {
    const size_t b = lds_yl<T, R1, R2>();
    if (b > limit) return hipErrorInvalidValue;
    hipError_t e = set_lds<T>((const void *)k_yl<T, R2, R1, false>, b);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_yl<T, R1, R2, true>, b);
    return e;
}
template <class T>
static hipError_t set_lds_yl(int py, size_t limit) {
    switch (py) {
        case 64: return set_lds_yl_shape<T, 8, 8>(limit);
        case 128: return set_lds_yl_shape<T, 8, 16>(limit);
        case 256: return set_lds_yl_shape<T, 16, 16>(limit);
        case 512: return set_lds_yl_shape<T, 16, 32>(limit);
        case 1024: return set_lds_yl_shape<T, 32, 32>(limit);
        case 2048: return set_lds_yl_shape<T, 64, 32>(limit);
        default: return hipErrorInvalidValue;
    }
}
template <class T, int R1, int R2>
static hipError_t set_lds_zf_shape(size_t limit) {
    const size_t b = lds_zf<T, R1, R2>();
    if (b > limit) return hipErrorInvalidValue;
    return set_lds<T>((const void *)k_zf<T, R1, R2, zf_units<R1, R2>()>, b);
}
template <class T>
static hipError_t set_lds_zf(int pz, size_t limit) {
    switch (pz) {
        case 4: return set_lds_zf_shape<T, 2, 2>(limit);
        case 8: return set_lds_zf_shape<T, 4, 2>(limit);
        case 16: return set_lds_zf_shape<T, 4, 4>(limit);
        case 32: return set_lds_zf_shape<T, 8, 4>(limit);
        case 64: return set_lds_zf_shape<T, 8, 8>(limit);
        case 128: return set_lds_zf_shape<T, 16, 8>(limit);
        case 256: return set_lds_zf_shape<T, 16, 16>(limit);
        default: return hipErrorInvalidValue;
    }
}

template <class T, int R1, int R2>
static hipError_t set_lds_xxp_shape(size_t bytes) {
    // the film's first and final x passes run with kXpfThreads too
    constexpr int TP = kXpfThreads, CF = xpf_chunk<R1, R2>();
    hipError_t e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 0, TP, CF>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 1, TP, CF>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 2, TP, CF>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 3, TP, CF>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 4, TP, CF>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_fwd_x<T, R1, R2, 5, TP, CF>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_inv_x<T, 0, R1, R2, TP, CF>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xxp<T, R1, R2, 1>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xxp<T, R1, R2, 2>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xxp<T, R1, R2, 3>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xxp<T, R1, R2, 4>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_xxp<T, R1, R2, 5>, bytes);
    return e;
}
template <class T>
static hipError_t set_lds_xxp(int xfast, size_t bytes) {
    switch (xfast) {
        case 512: return set_lds_xxp_shape<T, 16, 32>(bytes);
        case 1024: return set_lds_xxp_shape<T, 32, 32>(bytes);
        default: return hipErrorInvalidValue;
    }
}

template <class T>
const char *configure(Args<T> &A, size_t lds_limit, int rows_req, int xfast_req) {
    if (!A.demag) return "";
    A.xfast = 0;
    {
        int r1 = 0, r2 = 0;
        if (A.tw4x && yfast_shape(A.px, r1, r2) && xfast_req != 0) A.xfast = A.px;
    }
    auto fits_x = [&](int rows) {
        A.rows_per_wg = rows;
        A.nseq_x = (3 * rows + 1) / 2;
        // register four-step passes: the packed sequences sit ceil(32/nseq)
        // complex values further apart than px, so that the lanes of one kx
        // in the spectrum load/store phases (one per sequence) fall in
        // different LDS banks
        A.xsb = A.xfast ? A.px + (32 + A.nseq_x - 1) / A.nseq_x : A.px;
        return lds_x(A) <= lds_limit;
    };
    int rows = 2;
    if (rows_req > 0) {
        if (rows_req > kMaxRowsWg || !fits_x(rows_req)) return "requested x tile does not fit";
        rows = rows_req;
    } else if (A.nrows >= 4096 && fits_x(4)) {
        rows = 4;
    } else if (!fits_x(2)) {
        rows = 1;
    }
    if (!fits_x(rows)) return "padded x length does not fit the LDS of one work-group";
    A.yfast = 0;
    {
        int r1 = 0, r2 = 0;
        if (A.nz == 1 && A.tw4y && !A.gtens && yfast_shape(A.py, r1, r2)) A.yfast = A.py;
    }
    A.ylfast = 0;
    A.zfast = 0;
    if (A.nz > 1) {
        // register four-step passes where the shape has one and its LDS fits
        int r1 = 0, r2 = 0;
        if (A.tw4y && A.tw4y_f && yfast_shape(A.py, r1, r2) && set_lds_yl<T>(A.py, lds_limit) == hipSuccess)
            A.ylfast = A.py;
        if (A.tw4z && !A.gtens && zfast_shape(A.pz, r1, r2) && set_lds_zf<T>(A.pz, lds_limit) == hipSuccess)
            A.zfast = A.pz;
    }
    if (!A.yfast && !A.ylfast && lds_y(A) > lds_limit) return "padded y length does not fit the LDS of one work-group";
    if (A.nz > 1 && !A.zfast) {
        int ty = 16;
        while (ty > 1) {
            A.ty = ty;
            if (lds_z(A) <= lds_limit) break;
            ty >>= 1;
        }
        A.ty = ty;
        if (lds_z(A) > lds_limit) return "padded z length does not fit the LDS of one work-group";
    } else {
        A.ty = 1;
    }
    if (set_lds_x<T>(A.xfast, lds_x(A)) != hipSuccess) return "cannot raise LDS limit (x passes)";
    if (!A.yfast && set_lds<T>((const void *)k_y<T, 0>, lds_y(A)) != hipSuccess) return "cannot raise LDS limit (y)";
    if (A.yfast == 2048 && (lds_yf<T, 64, 32>() > lds_limit ||
                            set_lds<T>((const void *)k_yf<T, 64, 32>, lds_yf<T, 64, 32>()) != hipSuccess))
        return "cannot raise LDS limit (four-step y)";
    // (FP32 only: the FP64 form measured 2x slower over the whole film step)
    A.yf8 = sizeof(T) == 4 && A.yfast == 1024 && A.tw8 && A.yf8_allowed && lds_yf8<T>() <= lds_limit &&
            set_lds<T>((const void *)k_yf8<T>, lds_yf8<T>()) == hipSuccess;
    if (A.yfast == 1024 && set_lds<T>((const void *)k_yf<T, 32, 32>, lds_yf<T, 32, 32>()) != hipSuccess)
        return "cannot raise LDS limit (four-step y)";
    if (A.nz > 1 && !A.ylfast && set_lds<T>((const void *)k_y<T, 1>, lds_y(A)) != hipSuccess)
        return "cannot raise LDS limit (y)";
    if (A.nz > 1 && !A.ylfast && set_lds<T>((const void *)k_y<T, 2>, lds_y(A)) != hipSuccess)
        return "cannot raise LDS limit (y)";
    if (A.nz > 1 && !A.zfast && set_lds<T>((const void *)k_z<T>, lds_z(A)) != hipSuccess)
        return "cannot raise LDS limit (z)";
    // film joined x pass with its inputs prefetched into LDS (FP32; the
    // environment variable FULLMAG_FDM_XPASS_PREFETCH=0 keeps k_xx)
    A.xpf = 0;
    A.xpf_off = 0;
    if constexpr (sizeof(T) == 4) {
        const char *env = std::getenv("FULLMAG_FDM_XPASS_PREFETCH");
        const bool allowed = !(env && env[0] == '0');
        // px = 1024: the three-step FFTs transpose inside each sequence's
        // buffer (kF8Ta values); 32 more keeps the stride's bank offset
        const int xsb0 = A.xsb;
        if (A.xfast == 1024 && A.tw8) A.xsb = std::max(A.xsb, A.px + 32 + (32 + A.nseq_x - 1) / A.nseq_x);
        const int off = (int)((lds_x(A) + 15) / 16 * 16);
        const size_t bytes = (size_t)off + xpf_floats<T>(A.nx, kMaxK) * sizeof(T) + (size_t)kXpfRows * A.nx;
        if (allowed && A.nz == 1 && (A.xfast == 512 || (A.xfast == 1024 && A.tw8)) && A.rows_per_wg == kXpfRows &&
            A.nrows % kXpfRows == 0 && !A.region && !A.xt && A.method != M_ABM3 && bytes <= lds_limit &&
            set_lds_xxp<T>(A.xfast, bytes) == hipSuccess) {
            A.xpf = 1;
            A.xpf_off = off;
            // the other x-pass kernels see the wider stride too
            if (A.xsb != xsb0 && set_lds_x<T>(A.xfast, lds_x(A)) != hipSuccess)
                return "cannot raise LDS limit (x passes)";
        } else {
            A.xsb = xsb0;
        }
    }
    return "";
}


template <class T, int R1, int R2>
static void launch_yf_shape(const Args<T> &A, hipStream_t st) {
    constexpr int TS = R1 > R2 ? R1 : R2;
    k_yf<T, R1, R2><<<bgrid(A, dim3(A.nkx)), dim3(3 * TS), lds_yf<T, R1, R2>(), st>>>(A.self_d);
}

template <class T>
static void launch_yf(const Args<T> &A, hipStream_t st) {
    switch (A.yfast) {
        case 64: launch_yf_shape<T, 8, 8>(A, st); break;
        case 128: launch_yf_shape<T, 8, 16>(A, st); break;
        case 256: launch_yf_shape<T, 16, 16>(A, st); break;
        case 512: launch_yf_shape<T, 16, 32>(A, st); break;
        case 1024:
            if (A.yf8) k_yf8<T><<<bgrid(A, dim3(A.nkx)), dim3(384), lds_yf8<T>(), st>>>(A.self_d);
            else launch_yf_shape<T, 32, 32>(A, st);
            break;
        case 2048: launch_yf_shape<T, 64, 32>(A, st); break;
        default: break;
    }
}

template <class T>
long long num_partials(const Args<T> &A) {
    if (A.demag) return (A.nrows + A.rows_per_wg - 1) / A.rows_per_wg;
    return (A.ncell + kThreads - 1) / kThreads;
}

template <class T, int R1, int R2>
static void launch_fwd_x_shape(const Args<T> &A, int s, hipStream_t st) {
    const dim3 g((unsigned)num_partials(A)), b(kThreads);
    const size_t l = lds_x(A);
    if constexpr (xpf_shape<T, R1, R2>()) {
        if (A.xpf) {
            constexpr int TP = kXpfThreads, CF = xpf_chunk<R1, R2>();
            const dim3 bf(TP);
            switch (A.stage_nk[s]) {
                case 0: k_fwd_x<T, R1, R2, 0, TP, CF><<<bgrid(A, g), bf, l, st>>>(A.self_d, s); break;
                case 1: k_fwd_x<T, R1, R2, 1, TP, CF><<<bgrid(A, g), bf, l, st>>>(A.self_d, s); break;
                case 2: k_fwd_x<T, R1, R2, 2, TP, CF><<<bgrid(A, g), bf, l, st>>>(A.self_d, s); break;
                case 3: k_fwd_x<T, R1, R2, 3, TP, CF><<<bgrid(A, g), bf, l, st>>>(A.self_d, s); break;
                case 4: k_fwd_x<T, R1, R2, 4, TP, CF><<<bgrid(A, g), bf, l, st>>>(A.self_d, s); break;
                default: k_fwd_x<T, R1, R2, 5, TP, CF><<<bgrid(A, g), bf, l, st>>>(A.self_d, s); break;
            }
            return;
        }
    }
    switch (A.stage_nk[s]) {
        case 0: k_fwd_x<T, R1, R2, 0><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 1: k_fwd_x<T, R1, R2, 1><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 2: k_fwd_x<T, R1, R2, 2><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 3: k_fwd_x<T, R1, R2, 3><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 4: k_fwd_x<T, R1, R2, 4><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        default: k_fwd_x<T, R1, R2, 5><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
    }
}
// Kernel feature set of a plan: bit 0 region exchange table, bit 1 further
// terms (Args::xt).
template <class T>
static int feature_set(const Args<T> &A) {
    return (A.region ? 1 : 0) | (A.xt ? 2 : 0);
}
template <class T, int R1, int R2>
static void launch_inv_x_shape(const Args<T> &A, int s, int stats, hipStream_t st) {
    const dim3 g((unsigned)num_partials(A)), b(kThreads);
    if constexpr (xpf_shape<T, R1, R2>()) {
        if (A.xpf) {
            // Args::xpf: uniform exchange, no further terms (feature set 0)
            constexpr int TP = kXpfThreads, CF = xpf_chunk<R1, R2>();
            k_inv_x<T, 0, R1, R2, TP, CF><<<bgrid(A, g), dim3(TP), lds_x(A), st>>>(A.self_d, s, stats);
            return;
        }
    }
    switch (feature_set(A)) {
        case 0: k_inv_x<T, 0, R1, R2><<<bgrid(A, g), b, lds_x(A), st>>>(A.self_d, s, stats); break;
        case 1: k_inv_x<T, 1, R1, R2><<<bgrid(A, g), b, lds_x(A), st>>>(A.self_d, s, stats); break;
        case 2: k_inv_x<T, 2, R1, R2><<<bgrid(A, g), b, lds_x(A), st>>>(A.self_d, s, stats); break;
        default: k_inv_x<T, 3, R1, R2><<<bgrid(A, g), b, lds_x(A), st>>>(A.self_d, s, stats); break;
    }
}
template <class T, int FT, int R1, int R2>
static void launch_xx_ft(const Args<T> &A, int s, hipStream_t st) {
    const dim3 g((unsigned)num_partials(A)), b(kThreads);
    const size_t l = lds_x(A);
    switch (A.stage_nk[s + 1]) {
        case 1: k_xx<T, FT, R1, R2, 1><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 2: k_xx<T, FT, R1, R2, 2><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 3: k_xx<T, FT, R1, R2, 3><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 4: k_xx<T, FT, R1, R2, 4><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        default: k_xx<T, FT, R1, R2, 5><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
    }
}
template <class T, int R1, int R2>
static void launch_xx_shape(const Args<T> &A, int s, hipStream_t st) {
    switch (feature_set(A)) {
        case 0: launch_xx_ft<T, 0, R1, R2>(A, s, st); break;
        case 1: launch_xx_ft<T, 1, R1, R2>(A, s, st); break;
        case 2: launch_xx_ft<T, 2, R1, R2>(A, s, st); break;
        default: launch_xx_ft<T, 3, R1, R2>(A, s, st); break;
    }
}
template <class T, int R1, int R2>
static void launch_xxp_shape(const Args<T> &A, int s, hipStream_t st) {
    const dim3 g((unsigned)num_partials(A)), b(kXpfThreads);
    const size_t l = (size_t)A.xpf_off + xpf_floats<T>(A.nx, kMaxK) * sizeof(T) + (size_t)kXpfRows * A.nx;
    switch (A.stage_nk[s + 1]) {
        case 1: k_xxp<T, R1, R2, 1><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 2: k_xxp<T, R1, R2, 2><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 3: k_xxp<T, R1, R2, 3><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        case 4: k_xxp<T, R1, R2, 4><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
        default: k_xxp<T, R1, R2, 5><<<bgrid(A, g), b, l, st>>>(A.self_d, s); break;
    }
}
// Fused end of stage s and start of stage s+1 (2 <= s < nstage).
template <class T>
static void launch_xx(const Args<T> &A, int s, hipStream_t st) {
    if constexpr (sizeof(T) == 4) {
        if (A.xpf) {
            if (A.xfast == 512) launch_xxp_shape<T, 16, 32>(A, s, st);
            else launch_xxp_shape<T, 32, 32>(A, s, st);
            return;
        }
    }
    switch (A.xfast) {
        case 64: launch_xx_shape<T, 8, 8>(A, s, st); break;
        case 128: launch_xx_shape<T, 8, 16>(A, s, st); break;
        case 256: launch_xx_shape<T, 16, 16>(A, s, st); break;
        case 512: launch_xx_shape<T, 16, 32>(A, s, st); break;
        case 1024: launch_xx_shape<T, 32, 32>(A, s, st); break;
        case 2048: launch_xx_shape<T, 64, 32>(A, s, st); break;
        default: launch_xx_shape<T, 0, 0>(A, s, st); break;
    }
}
template <class T>
static void launch_fwd_x(const Args<T> &A, int s, hipStream_t st) {
    switch (A.xfast) {
        case 64: launch_fwd_x_shape<T, 8, 8>(A, s, st); break;
        case 128: launch_fwd_x_shape<T, 8, 16>(A, s, st); break;
        case 256: launch_fwd_x_shape<T, 16, 16>(A, s, st); break;
        case 512: launch_fwd_x_shape<T, 16, 32>(A, s, st); break;
        case 1024: launch_fwd_x_shape<T, 32, 32>(A, s, st); break;
        case 2048: launch_fwd_x_shape<T, 64, 32>(A, s, st); break;
        default: launch_fwd_x_shape<T, 0, 0>(A, s, st); break;
    }
}
template <class T>
static void launch_inv_x(const Args<T> &A, int s, int stats, hipStream_t st) {
    switch (A.xfast) {
        case 64: launch_inv_x_shape<T, 8, 8>(A, s, stats, st); break;
        case 128: launch_inv_x_shape<T, 8, 16>(A, s, stats, st); break;
        case 256: launch_inv_x_shape<T, 16, 16>(A, s, stats, st); break;
        case 512: launch_inv_x_shape<T, 16, 32>(A, s, stats, st); break;
        case 1024: launch_inv_x_shape<T, 32, 32>(A, s, stats, st); break;
        case 2048: launch_inv_x_shape<T, 64, 32>(A, s, stats, st); break;
        default: launch_inv_x_shape<T, 0, 0>(A, s, stats, st); break;
    }
}
template <class T, int R1, int R2>
static void launch_yl_shape(const Args<T> &A, bool inv, hipStream_t st) {
    constexpr int TS = R1 > R2 ? R1 : R2;
    const int nt = yl_threads_of<R1, R2>(A.yl_threads), lpb = nt / TS;
    const unsigned g = (unsigned)((3ll * A.nzs * A.nkx + lpb - 1) / lpb);
    if (inv) k_yl<T, R1, R2, true><<<bgrid(A, dim3(g)), dim3(nt), lds_yl<T, R1, R2>(nt), st>>>(A.self_d);
    else k_yl<T, R2, R1, false><<<bgrid(A, dim3(g)), dim3(nt), lds_yl<T, R1, R2>(nt), st>>>(A.self_d);
}
// 3-D y pass (forward: inv = false; inverse: inv = true).
template <class T>
static void launch_y3(const Args<T> &A, bool inv, hipStream_t st) {
    switch (A.ylfast) {
        case 64: launch_yl_shape<T, 8, 8>(A, inv, st); return;
        case 128: launch_yl_shape<T, 8, 16>(A, inv, st); return;
        case 256: launch_yl_shape<T, 16, 16>(A, inv, st); return;
        case 512: launch_yl_shape<T, 16, 32>(A, inv, st); return;
        case 1024: launch_yl_shape<T, 32, 32>(A, inv, st); return;
        case 2048: launch_yl_shape<T, 64, 32>(A, inv, st); return;
        default: break;
    }
    if (inv) k_y<T, 2><<<bgrid(A, dim3(A.nkx * A.nz)), dim3(kThreads), lds_y(A), st>>>(A.self_d);
    else k_y<T, 1><<<bgrid(A, dim3(A.nkx * A.nz)), dim3(kThreads), lds_y(A), st>>>(A.self_d);
}
template <class T, int R1, int R2>
static void launch_zf_shape(const Args<T> &A, hipStream_t st) {
    constexpr int TS = R1 > R2 ? R1 : R2;
    constexpr int UB = zf_units<R1, R2>();
    const int ub = zf_ub_of<R1, R2>(A.zf_ub);
    const unsigned g = (unsigned)(A.nkx * ((A.py + ub - 1) / ub));
    k_zf<T, R1, R2, UB><<<bgrid(A, dim3(g)), dim3(ub * TS), lds_zf<T, R1, R2>(ub), st>>>(A.self_d);
}
// 3-D z pass (forward FFT, tensor multiply, inverse FFT).
template <class T>
static void launch_z3(const Args<T> &A, hipStream_t st) {
    switch (A.zfast) {
        case 4: launch_zf_shape<T, 2, 2>(A, st); return;
        case 8: launch_zf_shape<T, 4, 2>(A, st); return;
        case 16: launch_zf_shape<T, 4, 4>(A, st); return;
        case 32: launch_zf_shape<T, 8, 4>(A, st); return;
        case 64: launch_zf_shape<T, 8, 8>(A, st); return;
        case 128: launch_zf_shape<T, 16, 8>(A, st); return;
        case 256: launch_zf_shape<T, 16, 16>(A, st); return;
        default: break;
    }
    const unsigned gz = (unsigned)(A.nkx * ((A.py + A.ty - 1) / A.ty));
    k_z<T><<<bgrid(A, dim3(gz)), dim3(kThreads), lds_z(A), st>>>(A.self_d);
}

// The y (and for 3-D z) passes between the forward and inverse x passes.
template <class T>
static void launch_yz(const Args<T> &A, hipStream_t st) {
    if (A.nz == 1) {
        if (A.yfast) launch_yf(A, st);
        else k_y<T, 0><<<bgrid(A, dim3(A.nkx)), dim3(kThreads), lds_y(A), st>>>(A.self_d);
    } else {
        launch_y3(A, false, st);
        if (A.lmix) {
            const unsigned g = (unsigned)(((long long)A.nkx * A.py + kThreads - 1) / kThreads);
            k_lmix<T><<<bgrid(A, dim3(g)), dim3(kThreads), 0, st>>>(A.self_d);
        } else {
            launch_z3(A, st);
        }
        launch_y3(A, true, st);
    }
}

// Local pass without demag (one cell per thread).
template <class T>
static void launch_local(const Args<T> &A, int s, int stats, hipStream_t st) {
    const dim3 g((unsigned)num_partials(A)), b(kThreads);
    switch (feature_set(A)) {
        case 0: k_local<T, 0><<<bgrid(A, g), b, 0, st>>>(A.self_d, s, stats); break;
        case 1: k_local<T, 1><<<bgrid(A, g), b, 0, st>>>(A.self_d, s, stats); break;
        case 2: k_local<T, 2><<<bgrid(A, g), b, 0, st>>>(A.self_d, s, stats); break;
        default: k_local<T, 3><<<bgrid(A, g), b, 0, st>>>(A.self_d, s, stats); break;
    }
}

// Closes (open = 0) or reopens (open = 1) the k1 evaluation of a thermal
// trial: bit 1 of Ctl::done makes every kernel return at once.  k1 is
// needed by DP45 only without a valid FSAL stage and by ABM3 only in its
// Heun start-up phase.
template <class T>
__global__ void k_k1_gate(const Args<T> *__restrict__ AP, int open) {
    if (threadIdx.x != 0) return;
    const Args<T> &A = AP[blockIdx.y];
    Ctl *c = A.ctl;
    if (open) {
        c->done &= ~2;
        return;
    }
    const bool need = A.method == M_DP45 ? c->k1_redo != 0 : A.method == M_ABM3 ? c->abm_count < 3 : true;
    if (!need) c->done |= 2;
}

// With a thermal field the RHS of a state depends on the step counter and
// on the dt of the attempt, so k1 is evaluated at the start of the attempt
// wherever the CPU engine's integrators evaluate it (fdm_integrators.rs):
// every attempt for Heun, RK4, RK23 and the ABM3 start-up, for DP45 the
// first attempt after create/upload and every retry (otherwise the FSAL
// stage of the previous step is k1).  Without a thermal field k1 is the
// previous step's last stage and nothing is launched.
template <class T>
hipError_t enqueue_eval(const Args<T> &A, int s, int control, hipStream_t st);
template <class T>
hipError_t enqueue_k1(const Args<T> &A, hipStream_t st) {
    if (!A.has_th) return hipSuccess;
    const bool gated = A.method == M_DP45 || A.method == M_ABM3;
    if (gated) k_k1_gate<T><<<bgrid(A, dim3(1)), dim3(64), 0, st>>>(A.self_d, 0);
    const hipError_t e = enqueue_eval(A, 1, 0, st);
    if (e != hipSuccess) return e;
    if (gated) k_k1_gate<T><<<bgrid(A, dim3(1)), dim3(64), 0, st>>>(A.self_d, 1);
    return hipGetLastError();
}

// One RHS evaluation of stage s (1 = the current state m0).  control:
// 0 none, 1 step controller, 2 statistics only (k_control after the final
// kernel; implies per-block statistics).
template <class T>
hipError_t enqueue_eval(const Args<T> &A, int s, int control, hipStream_t st) {
    const int stats = control != 0;
    const unsigned gpart = (unsigned)num_partials(A);
    if (A.demag) {
        launch_fwd_x(A, s, st);
        launch_yz(A, st);
        launch_inv_x(A, s, stats, st);
    } else {
        if (s >= 2) k_stage<T><<<bgrid(A, dim3(gpart)), dim3(kThreads), 0, st>>>(A.self_d, s);
        launch_local(A, s, stats, st);
    }
    if (control) k_control<T><<<bgrid(A, dim3(1)), dim3(kThreads), 0, st>>>(A.self_d, (int)gpart, control == 2);
    return hipGetLastError();
}

// Stages 2..nstage of one trial step and the controller.  With demag, the
// inverse x pass of stage s and the forward x pass of stage s+1 share one
// launch.
template <class T>
hipError_t enqueue_trial(const Args<T> &A, int nstage, hipStream_t st) {
    {
        const hipError_t e = enqueue_k1(A, st);
        if (e != hipSuccess) return e;
    }
    if (!A.demag) {
        for (int s = 2; s <= nstage; ++s) {
            const hipError_t e = enqueue_eval(A, s, s == nstage ? 1 : 0, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    launch_fwd_x(A, 2, st);
    for (int s = 2; s <= nstage; ++s) {
        launch_yz(A, st);
        if (s < nstage) launch_xx(A, s, st);
        else launch_inv_x(A, s, 1, st);
    }
    k_control<T><<<bgrid(A, dim3(1)), dim3(kThreads), 0, st>>>(A.self_d, (int)num_partials(A), 0);
    return hipGetLastError();
}

// ---- adjoint launches
// F: the argument record of the field cotangent's convolution (A itself, or
// the FP64 demag context of an FP32 system on the same grid).
template <class T, class TF, int R1, int R2>
static void launch_adj_x_shape(const Args<T> &A, const Args<TF> &F, const AdjArgs<T> &J, int j, bool inverse,
                               hipStream_t st) {
    const dim3 g((unsigned)num_partials(F)), b(kThreads);
    if (inverse) k_adj_bx<T, TF, R1, R2><<<bgrid(A, g), b, lds_x(F), st>>>(A.self_d, F.self_d, J, j);
    else k_adj_fx<T, TF, R1, R2><<<bgrid(A, g), b, lds_x(F), st>>>(A.self_d, F.self_d, J, j);
}
template <class T, class TF>
static void launch_adj_x(const Args<T> &A, const Args<TF> &F, const AdjArgs<T> &J, int j, bool inverse,
                         hipStream_t st) {
    switch (F.xfast) {
        case 64: launch_adj_x_shape<T, TF, 8, 8>(A, F, J, j, inverse, st); break;
        case 128: launch_adj_x_shape<T, TF, 8, 16>(A, F, J, j, inverse, st); break;
        case 256: launch_adj_x_shape<T, TF, 16, 16>(A, F, J, j, inverse, st); break;
        case 512: launch_adj_x_shape<T, TF, 16, 32>(A, F, J, j, inverse, st); break;
        case 1024: launch_adj_x_shape<T, TF, 32, 32>(A, F, J, j, inverse, st); break;
        case 2048: launch_adj_x_shape<T, TF, 64, 32>(A, F, J, j, inverse, st); break;
        default: launch_adj_x_shape<T, TF, 0, 0>(A, F, J, j, inverse, st); break;
    }
}
template <class T, class TF, int R1, int R2>
static hipError_t set_lds_adj_shape(size_t bytes) {
    hipError_t e = set_lds<T>((const void *)k_adj_fx<T, TF, R1, R2>, bytes);
    if (e == hipSuccess) e = set_lds<T>((const void *)k_adj_bx<T, TF, R1, R2>, bytes);
    return e;
}
template <class T, class TF>
static hipError_t set_lds_adj(int xfast, size_t bytes) {
    switch (xfast) {
        case 64: return set_lds_adj_shape<T, TF, 8, 8>(bytes);
        case 128: return set_lds_adj_shape<T, TF, 8, 16>(bytes);
        case 256: return set_lds_adj_shape<T, TF, 16, 16>(bytes);
        case 512: return set_lds_adj_shape<T, TF, 16, 32>(bytes);
        case 1024: return set_lds_adj_shape<T, TF, 32, 32>(bytes);
        case 2048: return set_lds_adj_shape<T, TF, 64, 32>(bytes);
        default: return set_lds_adj_shape<T, TF, 0, 0>(bytes);
    }
}

// y/z passes of a demag context (exported per precision, so that an FP32
// system's adjoint can run the FP64 context's passes).
template <class T>
hipError_t enqueue_yz(const Args<T> &A, hipStream_t st) {
    launch_yz(A, st);
    return hipGetLastError();
}
#ifdef FM_PRIMARY
// the FP64 passes come from the FP64 translation unit
extern template hipError_t enqueue_yz<double>(const Args<double> &, hipStream_t);
#endif

// Work-groups of the adjoint per-block partial arrays (3 doubles each);
// F as in launch_adj_x.
template <class T>
long long adj_blocks(const Args<T> &A) {
    const long long cells = (A.ncell + kThreads - 1) / kThreads;
    const long long xs = A.demag ? num_partials(A) : 0;
    return std::max(cells, xs);
}

// Reverse sweep of stage j (1 <= j < nstage): beta, its demag convolution,
// cotangent of y_j with the normalisation and accumulation.
template <class T>
hipError_t enqueue_adj_stage(const Args<T> &A, const AdjArgs<T> &J, int j, hipStream_t st) {
    const unsigned gc = (unsigned)((A.ncell + kThreads - 1) / kThreads);
    if (A.demag) {
        launch_adj_x(A, A, J, j, false, st);
        launch_yz(A, st);
        launch_adj_x(A, A, J, j, true, st);
    } else {
        k_adj_beta<T><<<bgrid(A, dim3(gc)), dim3(kThreads), 0, st>>>(A.self_d, J, j);
        k_adj_local<T><<<bgrid(A, dim3(gc)), dim3(kThreads), 0, st>>>(A.self_d, J, j);
    }
    return hipGetLastError();
}
#ifdef FM_PRIMARY
// The same for an FP32 system whose field cotangent is convolved in the FP64
// demag context F of its grid.
hipError_t enqueue_adj_stage_f64(const Args<float> &A, const Args<double> &F, const AdjArgs<float> &J, int j,
                                 hipStream_t st) {
    launch_adj_x(A, F, J, j, false, st);
    const hipError_t e = enqueue_yz<double>(F, st);
    if (e != hipSuccess) return e;
    launch_adj_x(A, F, J, j, true, st);
    return hipGetLastError();
}
hipError_t adj_configure_f64(const Args<double> &F) { return set_lds_adj<float, double>(F.xfast, lds_x(F)); }
#endif
template <class T>
hipError_t enqueue_adj_init(const Args<T> &A, const AdjArgs<T> &J, int nstage, hipStream_t st) {
    const unsigned gc = (unsigned)((A.ncell + kThreads - 1) / kThreads);
    k_adj_init<T><<<bgrid(A, dim3(gc)), dim3(kThreads), 0, st>>>(A.self_d, J, nstage);
    return hipGetLastError();
}
#ifdef FM_PRIMARY
hipError_t enqueue_adj_step(long long *step, long long value, int add, hipStream_t st) {
    k_adj_step<<<dim3(1), dim3(64), 0, st>>>(step, value, add);
    return hipGetLastError();
}
#endif
template <class T>
hipError_t enqueue_adj_objective(const Args<T> &A, const AdjArgs<T> &J, hipStream_t st) {
    const unsigned gc = (unsigned)std::min<long long>((A.ncell + kThreads - 1) / kThreads, 1024);
    k_adj_objective<T><<<bgrid(A, dim3(gc)), dim3(kThreads), 0, st>>>(A.self_d, J);
    return hipGetLastError();
}
template <class T>
hipError_t adj_configure(const Args<T> &A) {
    return A.demag ? set_lds_adj<T, T>(A.xfast, lds_x(A)) : hipSuccess;
}

template <class T>
hipError_t enqueue_control(const Args<T> &A, int observe, hipStream_t st) {
    k_control<T><<<bgrid(A, dim3(1)), dim3(kThreads), 0, st>>>(A.self_d, (int)num_partials(A), observe);
    return hipGetLastError();
}

#ifdef FM_PRIMARY
bool yfast_factors(int py, int *r1, int *r2) { return yfast_shape(py, *r1, *r2); }
bool zfast_factors(int pz, int *r1, int *r2) { return zfast_shape(pz, *r1, *r2); }
#endif

template hipError_t upload_methods<FM_REAL>();
template const char *configure<FM_REAL>(Args<FM_REAL> &, size_t, int, int);
template long long num_partials<FM_REAL>(const Args<FM_REAL> &);
template hipError_t enqueue_eval<FM_REAL>(const Args<FM_REAL> &, int, int, hipStream_t);
template hipError_t enqueue_trial<FM_REAL>(const Args<FM_REAL> &, int, hipStream_t);
template hipError_t enqueue_k1<FM_REAL>(const Args<FM_REAL> &, hipStream_t);
template hipError_t enqueue_control<FM_REAL>(const Args<FM_REAL> &, int, hipStream_t);
template long long adj_blocks<FM_REAL>(const Args<FM_REAL> &);
template hipError_t enqueue_adj_stage<FM_REAL>(const Args<FM_REAL> &, const AdjArgs<FM_REAL> &, int, hipStream_t);
template hipError_t enqueue_adj_init<FM_REAL>(const Args<FM_REAL> &, const AdjArgs<FM_REAL> &, int, hipStream_t);
template hipError_t enqueue_adj_objective<FM_REAL>(const Args<FM_REAL> &, const AdjArgs<FM_REAL> &, hipStream_t);
template hipError_t adj_configure<FM_REAL>(const Args<FM_REAL> &);
template hipError_t enqueue_yz<FM_REAL>(const Args<FM_REAL> &, hipStream_t);

}  // namespace fm

namespace fm {

// One launch of a single kernel kind (for event timing): 0 x-forward or
// stage combine, 1 y pass, 2 z pass, 3 y inverse, 4 x-inverse/local pass,
// 5 controller, 6 fused x-inverse/local of stage s + x-forward of s+1.  Returns hipErrorInvalidValue for a kind the grid lacks.
template <class T>
hipError_t enqueue_kind(const Args<T> &A, int kind, int s, hipStream_t st) {
    const unsigned gpart = (unsigned)num_partials(A);
    const bool film = A.nz == 1;
    switch (kind) {
        case 0:
            if (A.demag) launch_fwd_x(A, s, st);
            else k_stage<T><<<bgrid(A, dim3(gpart)), dim3(kThreads), 0, st>>>(A.self_d, s);
            break;
        case 1:
            if (!A.demag) return hipErrorInvalidValue;
            if (film && A.yfast) launch_yf(A, st);
            else if (film) k_y<T, 0><<<bgrid(A, dim3(A.nkx)), dim3(kThreads), lds_y(A), st>>>(A.self_d);
            else launch_y3(A, false, st);
            break;
        case 2:
            if (!A.demag || film) return hipErrorInvalidValue;
            launch_z3(A, st);
            break;
        case 3:
            if (!A.demag || film) return hipErrorInvalidValue;
            launch_y3(A, true, st);
            break;
        case 4:
            if (A.demag) {
                launch_inv_x(A, s, 1, st);
            } else {
                launch_local(A, s, 1, st);
            }
            break;
        case 5:
            k_control<T><<<bgrid(A, dim3(1)), dim3(kThreads), 0, st>>>(A.self_d, (int)gpart, 0);
            break;
        case 6:
            if (!A.demag) return hipErrorInvalidValue;
            launch_xx(A, s, st);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t enqueue_kind<FM_REAL>(const Args<FM_REAL> &, int, int, hipStream_t);

// Preview bins of a SoA field (reference downsample_field_preview_kernel,
// native/backends/fdm/src/context.cu:254-317): preview cell (px, py, pz)
// averages the box x in [px fx/pnx, (px+1) fx/pnx), y likewise, z in
// [z0 + pz zs, z0 + (pz+1) zs), each range at least one cell and clipped to
// the grid; FP64 sums in z, y, x order; AoS output [x0,y0,z0, x1,...].  One
// thread per preview cell.
template <class T, class O>
__global__ void __launch_bounds__(kThreads) k_preview(const T *__restrict__ f, long long n, PreviewGeom g,
                                                      O *__restrict__ out) {
    const unsigned long long cnt = (unsigned long long)g.pnx * g.pny * g.pnz;
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const uint32_t px = (uint32_t)(i % g.pnx), py = (uint32_t)((i / g.pnx) % g.pny),
                   pz = (uint32_t)(i / ((unsigned long long)g.pnx * g.pny));
    auto range = [](uint32_t p, uint32_t full, uint32_t np, uint32_t &b, uint32_t &e) {
        b = (uint32_t)(((unsigned long long)p * full) / np);
        e = (uint32_t)(((unsigned long long)(p + 1) * full) / np);
        if (e <= b) e = b + 1;
        if (e > full) e = full;
    };
    uint32_t xb, xe, yb, ye;
    range(px, g.fx, g.pnx, xb, xe);
    range(py, g.fy, g.pny, yb, ye);
    uint32_t zb = g.z0 + pz * g.zs;
    if (zb >= g.fz) zb = g.fz - 1;
    uint32_t ze = g.z0 + (pz + 1) * g.zs;
    if (ze <= zb) ze = zb + 1;
    if (ze > g.fz) ze = g.fz;
    const gp<const T> F = G(f);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, c = 0.0;
    for (uint32_t z = zb; z < ze; ++z)
        for (uint32_t y = yb; y < ye; ++y)
            for (uint32_t x = xb; x < xe; ++x) {
                const long long q = ((long long)z * g.fy + y) * g.fx + x;
                a0 += (double)F[q];
                a1 += (double)F[n + q];
                a2 += (double)F[2 * n + q];
                c += 1.0;
            }
    const gp<O> o = G(out);
    o[3 * i] = (O)(a0 / c);
    o[3 * i + 1] = (O)(a1 / c);
    o[3 * i + 2] = (O)(a2 / c);
}

template <class T>
hipError_t enqueue_preview(const T *field, long long ncell, const PreviewGeom &g, void *out, int out_fp64,
                           hipStream_t st) {
    const unsigned long long cnt = (unsigned long long)g.pnx * g.pny * g.pnz;
    if (cnt == 0 || g.fx == 0 || g.fy == 0 || g.fz == 0) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((cnt + kThreads - 1) / kThreads));
    if (out_fp64) k_preview<T, double><<<grid, dim3(kThreads), 0, st>>>(field, ncell, g, static_cast<double *>(out));
    else k_preview<T, float><<<grid, dim3(kThreads), 0, st>>>(field, ncell, g, static_cast<float *>(out));
    return hipGetLastError();
}

template hipError_t enqueue_preview<FM_REAL>(const FM_REAL *, long long, const PreviewGeom &, void *, int, hipStream_t);

#ifdef FM_PRIMARY
// Staging array of the allocation-free preview entry point.  Each GPU holds
// its own copy until the process ends.  In host code the name `scratch` is
// only a stand-in: the device address of the current GPU's copy comes from
// hipGetSymbolAddress, looked up once per device and kept.
static __device__ __align__(256) unsigned char scratch[kPreviewScratchBytes];

hipError_t preview_scratch(int device, void **ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        // the device must be current: the symbol resolves to its copy
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}
#endif

}  // namespace fm
