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
#include <cmath>
#include <cstdio>
#include "engine.hpp"

namespace fm {

template <class T>
struct C2 {
    T x, y;
};

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

__constant__ Method c_methods[M_COUNT];

extern __shared__ __align__(16) unsigned char fm_smem[];

// ------------------------------------------------------------ butterflies --
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

// Generic odd prime radix (R <= kMaxRadix), O(R^2) DFT per butterfly.
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
        const C2<T> *s = src + sq * n;
        C2<T> v[kMaxRadix];
        for (int r = 0; r < R; ++r) {
            const int idx = j + r * nr;
            v[r] = (first && idx >= n_in) ? C2<T>{T(0), T(0)} : s[idx];
            if (ns > 1 && r > 0) {
                C2<T> w = tw[k * r * tstep];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        const int base = (j / ns) * ns * R + k;
        C2<T> *d = dst + sq * n;
        for (int o = 0; o < R; ++o) {
            C2<T> acc = {T(0), T(0)};
            for (int q = 0; q < R; ++q) {
                C2<T> w = tw[((q * o) % R) * wstep];
                if (INV) w.y = -w.y;
                acc = cadd(acc, cmul(v[q], w));
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
__device__ __forceinline__ int kslot(const Method &M, int j, int p) {
    return j == 1 ? p : (j == M.nstage ? 1 - p : j);
}

template <class T>
__device__ __forceinline__ const T *stage_input(const Args<T> &A, const Method &M, int s, int p) {
    const long long n3 = 3 * A.ncell;
    if (s <= 1) return A.m_all + p * n3;
    if (s == M.nstage) return A.m_all + (1 - p) * n3;
    return A.ybuf;
}

template <class T>
__device__ __forceinline__ T *stage_rhs(const Args<T> &A, const Method &M, int s, int p) {
    const int slot = s <= 1 ? p : kslot(M, s, p);
    return A.k_all + slot * 3 * A.ncell;
}

// y_s at `cell` (s >= 2 combines and stores; s <= 1 reads m0).
template <class T>
__device__ __forceinline__ void stage_state(const Args<T> &A, const Method &M, int s, int p, double dt, long long cell,
                                            T v[3]) {
    const long long N = A.ncell;
    const T *m0 = A.m_all + p * 3 * N;
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
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int q = 0; q < R.nk; ++q) {
        const T *k = A.k_all + kslot(M, R.kidx[q], p) * 3 * N;
        const double w = R.a[q];
        a0 += w * (double)k[cell];
        a1 += w * (double)k[N + cell];
        a2 += w * (double)k[2 * N + cell];
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
    T *out = s == M.nstage ? A.m_all + (1 - p) * 3 * N : A.ybuf;
    out[cell] = (T)x;
    out[N + cell] = (T)y;
    out[2 * N + cell] = (T)z;
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

// Block-wide reduction of Acc into out[0..kStatSlots) by thread 0 (fixed
// order, so deterministic).  Must be called by all threads.
__device__ void block_reduce_store(Acc &a, double *out) {
    __shared__ double red[kStatSlots][kThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < kStatSlots; ++i) {
        double x = a.v[i];
        for (int off = 32; off > 0; off >>= 1) {
            const double y = __shfl_xor(x, off, 64);
            x = slot_is_max(i) ? fmax(x, y) : x + y;
        }
        if (lane == 0) red[i][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < kStatSlots) {
        const int i = threadIdx.x;
        double x = red[i][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) x = slot_is_max(i) ? fmax(x, red[i][w]) : x + red[i][w];
        out[i] = x;
    }
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

// Evaluation time of stage s of the trial step from the control block.
template <class T>
__device__ __forceinline__ double stage_time(const Args<T> &A, const Method &M, int s, const Ctl *ctl, double dt) {
    return s <= 1 ? ctl->t : ctl->t + M.node[s] * dt;
}

// Local field terms, LLG torque, and (optionally) statistics for one cell.
template <class T, bool LUT>
__device__ __forceinline__ void local_cell(const Args<T> &A, const Method &M, int s, int p, double dt, long long cell,
                                           int x, int y, int z, const T *__restrict__ yin, T *__restrict__ kout,
                                           const T hd[3], T wf, bool stats, bool err, Acc &acc) {
    const long long N = A.ncell;
    const bool act = A.mask ? A.mask[cell] != 0 : true;
    if (!act) {
        kout[cell] = T(0);
        kout[N + cell] = T(0);
        kout[2 * N + cell] = T(0);
        if (A.obs_hex) {
            A.obs_hex[cell] = A.obs_hex[N + cell] = A.obs_hex[2 * N + cell] = T(0);
            A.obs_hd[cell] = A.obs_hd[N + cell] = A.obs_hd[2 * N + cell] = T(0);
            A.obs_heff[cell] = A.obs_heff[N + cell] = A.obs_heff[2 * N + cell] = T(0);
        }
        return;
    }
    const T m[3] = {yin[cell], yin[N + cell], yin[2 * N + cell]};
    const T ms_i = A.ms_cell ? A.ms_cell[cell] : A.ms;
    // neighbours: Neumann clamp at the box, inactive neighbour -> centre
    const long long sx = 1, sy = A.nx, sz = (long long)A.nx * A.ny;
    long long nb[6] = {x > 0 ? cell - sx : cell, x < A.nx - 1 ? cell + sx : cell,
                       y > 0 ? cell - sy : cell, y < A.ny - 1 ? cell + sy : cell,
                       z > 0 ? cell - sz : cell, z < A.nz - 1 ? cell + sz : cell};
    if (A.mask) {
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (A.mask[nb[q]] == 0) nb[q] = cell;
    }
    T hex[3];
    if (!LUT) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T *f = yin + c * N;
            const T mc = m[c];
            const T lap = (f[nb[1]] - T(2) * mc + f[nb[0]]) * A.idx2 + (f[nb[3]] - T(2) * mc + f[nb[2]]) * A.idy2 +
                          (f[nb[5]] - T(2) * mc + f[nb[4]]) * A.idz2;
            hex[c] = (A.ms_cell ? A.two_a_mu0 / ms_i : A.ex_pre) * lap;
        }
    } else {
        const uint32_t rc = A.region[cell];
        T aij[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) aij[q] = A.lut[rc * 256u + A.region[nb[q]]];
        const T w[6] = {A.idx2, A.idx2, A.idy2, A.idy2, A.idz2, A.idz2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T *f = yin + c * N;
            T acc2 = T(0);
#pragma unroll
            for (int q = 0; q < 6; ++q) acc2 += aij[q] * (f[nb[q]] - m[c]) * w[q];
            hex[c] = (A.ms_cell ? A.two_mu0inv / ms_i : A.lut_pre) * acc2;
        }
    }
    T hext[3] = {T(0), T(0), T(0)};
    if (A.has_ext) {
        hext[0] = A.hext[0];
        hext[1] = A.hext[1];
        hext[2] = A.hext[2];
    }
    if (A.wave_kind && (!A.wave_mask || A.wave_mask[cell])) {
        hext[0] += wf * T(A.wave_amp[0]);
        hext[1] += wf * T(A.wave_amp[1]);
        hext[2] += wf * T(A.wave_amp[2]);
    }
    T hani[3] = {T(0), T(0), T(0)};
    if (A.has_uni) {
        T u[3] = {A.axis[0], A.axis[1], A.axis[2]};
        if (A.axis_cell) {
            u[0] = A.axis_cell[cell];
            u[1] = A.axis_cell[N + cell];
            u[2] = A.axis_cell[2 * N + cell];
        }
        T c1 = A.uni_c1, c2 = A.uni_c2;
        if (A.ku1_cell || A.ku2_cell || A.ms_cell) {
            const T k1 = A.ku1_cell ? A.ku1_cell[cell] : A.ku1_u;
            const T k2 = A.ku2_cell ? A.ku2_cell[cell] : A.ku2_u;
            c1 = A.two_mu0inv * k1 / ms_i;
            c2 = T(2) * A.two_mu0inv * k2 / ms_i;
        }
        const T mu = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
        const T coef = c1 * mu + c2 * mu * mu * mu;
        hani[0] = u[0] * coef;
        hani[1] = u[1] * coef;
        hani[2] = u[2] * coef;
    }
    T h[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) h[c] = hex[c] + hd[c] + hext[c] + hani[c];
    // LLG: k = -gb (m x H + alpha m x (m x H))
    const T pr[3] = {m[1] * h[2] - m[2] * h[1], m[2] * h[0] - m[0] * h[2], m[0] * h[1] - m[1] * h[0]};
    const T dm[3] = {m[1] * pr[2] - m[2] * pr[1], m[2] * pr[0] - m[0] * pr[2], m[0] * pr[1] - m[1] * pr[0]};
    T k[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) k[c] = -A.gamma_bar * ((A.precession ? pr[c] : T(0)) + A.alpha * dm[c]);
    kout[cell] = k[0];
    kout[N + cell] = k[1];
    kout[2 * N + cell] = k[2];
    if (A.obs_hex) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            A.obs_hex[c * N + cell] = hex[c];
            A.obs_hd[c * N + cell] = hd[c];
            A.obs_heff[c * N + cell] = h[c];
        }
    }
    if (stats) {
        const double md[3] = {(double)m[0], (double)m[1], (double)m[2]};
        auto dotd = [&](const T *f) { return md[0] * (double)f[0] + md[1] * (double)f[1] + md[2] * (double)f[2]; };
        auto sq = [](const T *f) {
            return (double)f[0] * (double)f[0] + (double)f[1] * (double)f[1] + (double)f[2] * (double)f[2];
        };
        const double ep = A.ms_cell ? A.mu0_vol * (double)ms_i : A.e_pre;
        acc.v[ST_EEX] += -0.5 * ep * dotd(hex);
        acc.v[ST_EDEMAG] += -0.5 * ep * dotd(hd);
        acc.v[ST_EEXT] += -1.0 * ep * dotd(hext);
        acc.v[ST_EANI] += -0.5 * ep * dotd(hani);
        acc.v[ST_HEFF2] = fmax(acc.v[ST_HEFF2], sq(h));
        acc.v[ST_HD2] = fmax(acc.v[ST_HD2], sq(hd));
        acc.v[ST_RHS2] = fmax(acc.v[ST_RHS2], sq(k));
    }
    if (err) {
        double e0 = 0.0, e1 = 0.0, e2 = 0.0;
        for (int q = 0; q < M.nerr; ++q) {
            const int j = M.eidx[q];
            const double w = M.e[q];
            if (j == M.nstage) {
                e0 += w * (double)k[0];
                e1 += w * (double)k[1];
                e2 += w * (double)k[2];
            } else {
                const T *kk = A.k_all + kslot(M, j, p) * 3 * N;
                e0 += w * (double)kk[cell];
                e1 += w * (double)kk[N + cell];
                e2 += w * (double)kk[2 * N + cell];
            }
        }
        e0 *= dt;
        e1 *= dt;
        e2 *= dt;
        double v = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
        if (A.rtol > 0.0) {
            const T *m0 = A.m_all + p * 3 * N;
            const double a = (double)m0[cell], b = (double)m0[N + cell], c = (double)m0[2 * N + cell];
            v /= (A.atol + A.rtol * fmax(sqrt(a * a + b * b + c * c), 1e-30));
        }
        acc.v[ST_ERR] = fmax(acc.v[ST_ERR], v);
    }
}

// ---------------------------------------------------------------- kernels --
template <class T>
__global__ void __launch_bounds__(kThreads) k_fwd_x(Args<T> A, int s) {
    Ctl *ctl = A.ctl;
    if (ctl->done) return;
    const int p = ctl->parity;
    const double dt = ctl->dt;
    const Method &M = c_methods[A.method];
    C2<T> *b0 = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *b1 = b0 + A.nseq_x * A.px;
    C2<T> *tw = b1 + A.nseq_x * A.px;
    const int rows = A.rows_per_wg, nx = A.nx, px = A.px, nl = 3 * rows;
    const long long r0 = (long long)blockIdx.x * rows;
    load_twiddles<T>(tw, A.twx, px);
    for (int i = threadIdx.x; i < rows * nx; i += blockDim.x) {
        const int rr = i / nx, x = i - rr * nx;
        const long long row = r0 + rr;
        T v[3] = {T(0), T(0), T(0)};
        if (row < A.nrows) stage_state(A, M, s, p, dt, row * nx + x, v);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int l = 3 * rr + c;
            T *e = reinterpret_cast<T *>(&b0[(l >> 1) * px + x]);
            e[l & 1] = (A.ms_cell && row < A.nrows ? A.ms_cell[row * nx + x] : A.ms) * v[c];
        }
        if ((nl & 1) && rr == rows - 1) b0[(nl >> 1) * px + x].y = T(0);
    }
    __syncthreads();
    const C2<T> *res = lds_fft<T, false>(b0, b1, A.fx, tw, A.nseq_x, nx, px);
    C2<T> *S = reinterpret_cast<C2<T> *>(A.spec);
    const int nkx = A.nkx;
    for (int i = threadIdx.x; i < nl * nkx; i += blockDim.x) {
        const int rr = i % rows;
        const int rest = i / rows;
        const int kx = rest % nkx, c = rest / nkx;
        const long long row = r0 + rr;
        if (row >= A.nrows) continue;
        const int l = 3 * rr + c;
        const C2<T> *zq = res + (l >> 1) * px;
        const C2<T> zk = zq[kx], zm = zq[kx == 0 ? 0 : px - kx];
        C2<T> X;
        if (!(l & 1)) X = {(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
        else X = {(zk.y + zm.y) * T(0.5), (zm.x - zk.x) * T(0.5)};
        const long long yy = row % A.ny, zz = row / A.ny;
        S[(((long long)c * A.nzs + zz) * nkx + kx) * A.ly + yy] = X;
    }
}

// MODE 0: thin-film fused forward FFT + tensor multiply + inverse FFT.
// MODE 1: 3-D forward y FFT.   MODE 2: 3-D inverse y FFT.
template <class T, int MODE>
__global__ void __launch_bounds__(kThreads) k_y(Args<T> A) {
    if (A.ctl->done) return;
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

template <class T>
__global__ void __launch_bounds__(kThreads) k_z(Args<T> A) {
    if (A.ctl->done) return;
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
    const long long cs = (long long)nkx * (ny + 1) * (nz + 1);
    for (int i = threadIdx.x; i < TY * pz; i += blockDim.x) {
        const int t = i / pz, kz = i - t * pz;
        const int ky = ky0 + t;
        if (ky >= py) continue;
        const bool ylo = ky <= ny, zlo = kz <= nz;
        const int kyr = ylo ? ky : py - ky, kzr = zlo ? kz : pz - kz;
        const T sy = ylo ? T(1) : T(-1), sz = zlo ? T(1) : T(-1);
        const T *N = A.ntens + ((long long)kx * (ny + 1) + kyr) * (nz + 1) + kzr;
        const T nxx = N[0], nyy = N[cs], nzz = N[2 * cs];
        const T nxy = sy * N[3 * cs], nxz = sz * N[4 * cs], nyz = sy * sz * N[5 * cs];
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

template <class T, bool LUT>
__global__ void __launch_bounds__(kThreads) k_inv_x(Args<T> A, int s, int stats, int control) {
    Ctl *ctl = A.ctl;
    if (ctl->done) return;
    const int p = ctl->parity;
    const double dt = ctl->dt;
    const Method &M = c_methods[A.method];
    C2<T> *b0 = reinterpret_cast<C2<T> *>(fm_smem);
    C2<T> *b1 = b0 + A.nseq_x * A.px;
    C2<T> *tw = b1 + A.nseq_x * A.px;
    const int rows = A.rows_per_wg, nx = A.nx, px = A.px, nl = 3 * rows, nkx = A.nkx;
    const long long r0 = (long long)blockIdx.x * rows;
    load_twiddles<T>(tw, A.twx, px);
    const C2<T> *S = reinterpret_cast<const C2<T> *>(A.spec);
    auto line = [&](int l, int kk) -> C2<T> {
        if (l >= nl) return C2<T>{T(0), T(0)};
        const int rr = l / 3, c = l - 3 * (l / 3);
        const long long row = r0 + rr;
        if (row >= A.nrows) return C2<T>{T(0), T(0)};
        const long long yy = row % A.ny, zz = row / A.ny;
        return S[(((long long)c * A.nzs + zz) * nkx + kk) * A.ly + yy];
    };
    for (int i = threadIdx.x; i < A.nseq_x * px; i += blockDim.x) {
        const int sq = i / px, k = i - sq * px;
        const bool cj = k > nx;
        const int kk = cj ? px - k : k;
        C2<T> a = line(2 * sq, kk), b = line(2 * sq + 1, kk);
        if (cj) {
            a.y = -a.y;
            b.y = -b.y;
        }
        b0[i] = {a.x - b.y, a.y + b.x};
    }
    __syncthreads();
    const C2<T> *res = lds_fft<T, true>(b0, b1, A.fx, tw, A.nseq_x, px, nx);
    const T *yin = stage_input(A, M, s, p);
    T *kout = stage_rhs(A, M, s, p);
    const bool err = M.adaptive && s == M.nstage;
    const T scale = T(A.inv_p);
    const T wf = A.wave_kind ? wave_factor(A, stage_time(A, M, s, ctl, dt)) : T(0);
    Acc acc;
    acc.init();
    for (int i = threadIdx.x; i < rows * nx; i += blockDim.x) {
        const int rr = i / nx, x = i - rr * nx;
        const long long row = r0 + rr;
        if (row >= A.nrows) continue;
        T hd[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int l = 3 * rr + c;
            const C2<T> zv = res[(l >> 1) * px + x];
            hd[c] = ((l & 1) ? zv.y : zv.x) * scale;
        }
        const int yy = (int)(row % A.ny), zz = (int)(row / A.ny);
        local_cell<T, LUT>(A, M, s, p, dt, row * nx + x, x, yy, zz, yin, kout, hd, wf, stats != 0, err, acc);
    }
    if (stats || err) block_reduce_store(acc, A.partials + (long long)blockIdx.x * kStatSlots);
    if (control) finish_last_block(A, control);
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_stage(Args<T> A, int s) {
    Ctl *ctl = A.ctl;
    if (ctl->done) return;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= A.ncell) return;
    const Method &M = c_methods[A.method];
    T v[3];
    stage_state(A, M, s, ctl->parity, ctl->dt, cell, v);
}

template <class T, bool LUT>
__global__ void __launch_bounds__(kThreads) k_local(Args<T> A, int s, int stats, int control) {
    Ctl *ctl = A.ctl;
    if (ctl->done) return;
    const int p = ctl->parity;
    const double dt = ctl->dt;
    const Method &M = c_methods[A.method];
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool err = M.adaptive && s == M.nstage;
    Acc acc;
    acc.init();
    if (cell < A.ncell) {
        const int x = (int)(cell % A.nx);
        const int y = (int)((cell / A.nx) % A.ny);
        const int z = (int)(cell / ((long long)A.nx * A.ny));
        const T hd[3] = {T(0), T(0), T(0)};
        const T wf = A.wave_kind ? wave_factor(A, stage_time(A, M, s, ctl, dt)) : T(0);
        local_cell<T, LUT>(A, M, s, p, dt, cell, x, y, z, stage_input(A, M, s, p), stage_rhs(A, M, s, p), hd, wf,
                           stats != 0, err, acc);
    }
    if (stats || err) block_reduce_store(acc, A.partials + (long long)blockIdx.x * kStatSlots);
    if (control) finish_last_block(A, control);
}

// Reduction of all per-block partials and the step controller, executed by
// one work-group.  observe != 0: only store the statistics.
template <class T>
__device__ void control_block(const Args<T> &A, int nparts, int observe) {
    Ctl *c = A.ctl;
    Acc acc;
    acc.init();
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        const double *q = A.partials + (long long)i * kStatSlots;
#pragma unroll
        for (int k = 0; k < kStatSlots; ++k) {
            // agent-scope loads: the partials were written by other CUs
            const double v = __hip_atomic_load(q + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            acc.v[k] = slot_is_max(k) ? fmax(acc.v[k], v) : acc.v[k] + v;
        }
    }
    __shared__ double tot[kStatSlots];
    block_reduce_store(acc, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const Method &M = c_methods[A.method];
    if (observe) {
        for (int k = 0; k < kStatSlots; ++k) c->stats[k] = tot[k];
        return;
    }
    const double dt = c->dt;
    c->attempts += 1;
    bool accept = true;
    if (M.adaptive) {
        const double err = tot[ST_ERR];
        c->err = err;
        const double thr = A.rtol > 0.0 ? 1.0 : A.atol;
        accept = err <= thr || dt <= A.dt_min;
        const double ratio =
            fmax(fmin(A.headroom * pow(thr / (accept ? fmax(err, 1e-30) : err), 0.2), A.growth), A.shrink);
        const double dtn = fmin(fmax(dt * ratio, A.dt_min), A.dt_max);
        if (accept) c->dt_next = dtn;
        else c->dt = dtn;
    } else {
        c->dt_next = 0.0;
    }
    if (!accept) {
        c->rejected += 1;
        c->accepted = 0;
        return;
    }
    for (int k = 0; k < kStatSlots; ++k) c->stats[k] = tot[k];
    c->t += dt;
    c->step += 1;
    c->parity ^= 1;
    c->dt_used = dt;
    c->accepted = 1;
    if (c->loop) {
        if (M.adaptive) c->dt_carry = c->dt_next;
        if (c->t >= c->t_end || c->step >= c->step_end) {
            c->done = 1;
        } else {
            double d = fmin(c->dt_carry, c->t_end - c->t);
            if (M.adaptive) d = fmin(fmax(d, A.dt_min), A.dt_max);
            c->dt = d;
        }
    }
}

// Stand-alone controller launch (used for timing and by callers that keep
// the partial-producing kernel and the controller apart).
template <class T>
__global__ void __launch_bounds__(kThreads) k_control(Args<T> A, int nparts, int observe) {
    if (A.ctl->done) return;
    control_block(A, nparts, observe);
}

// End of a partial-producing kernel: the last work-group to finish reduces
// the partials of the whole grid and runs the controller (control 1) or
// stores the statistics (control 2), then re-arms the counter for the next
// launch.  Every work-group of the grid must reach this point.
template <class T>
__device__ void finish_last_block(const Args<T> &A, int control) {
    __shared__ int is_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned prev = atomicAdd(A.counter, 1u);
        is_last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    control_block(A, (int)gridDim.x, control == 2);
    if (threadIdx.x == 0) atomicExch(A.counter, 0u);
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
}

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

int method_nstage(int method) {
    Method tab[M_COUNT];
    build_methods(tab);
    return tab[method].nstage;
}

template <class T>
static size_t lds_x(const Args<T> &A) { return (2ull * A.nseq_x + 1) * A.px * sizeof(C2<T>); }
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
template <class T>
const char *configure(Args<T> &A, size_t lds_limit) {
    if (!A.demag) return "";
    auto fits_x = [&](int rows) {
        A.rows_per_wg = rows;
        A.nseq_x = (3 * rows + 1) / 2;
        return lds_x(A) <= lds_limit;
    };
    int rows = 2;
    if (A.nrows >= 4096 && fits_x(4)) rows = 4;
    else if (!fits_x(2)) rows = 1;
    if (!fits_x(rows)) return "padded x length does not fit the LDS of one work-group";
    if (lds_y(A) > lds_limit) return "padded y length does not fit the LDS of one work-group";
    if (A.nz > 1) {
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
    if (set_lds<T>((const void *)k_fwd_x<T>, lds_x(A)) != hipSuccess) return "cannot raise LDS limit (x fwd)";
    if (set_lds<T>((const void *)k_inv_x<T, false>, lds_x(A)) != hipSuccess) return "cannot raise LDS limit (x inv)";
    if (set_lds<T>((const void *)k_inv_x<T, true>, lds_x(A)) != hipSuccess) return "cannot raise LDS limit (x inv)";
    if (set_lds<T>((const void *)k_y<T, 0>, lds_y(A)) != hipSuccess) return "cannot raise LDS limit (y)";
    if (set_lds<T>((const void *)k_y<T, 1>, lds_y(A)) != hipSuccess) return "cannot raise LDS limit (y)";
    if (set_lds<T>((const void *)k_y<T, 2>, lds_y(A)) != hipSuccess) return "cannot raise LDS limit (y)";
    if (A.nz > 1 && set_lds<T>((const void *)k_z<T>, lds_z(A)) != hipSuccess) return "cannot raise LDS limit (z)";
    return "";
}

template <class T>
long long num_partials(const Args<T> &A) {
    if (A.demag) return (A.nrows + A.rows_per_wg - 1) / A.rows_per_wg;
    return (A.ncell + kThreads - 1) / kThreads;
}

// One RHS evaluation of stage s (1 = the current state m0).  control:
// 0 none, 1 step controller, 2 statistics only (fused into the last
// work-group of the final kernel; implies per-block statistics).
template <class T>
hipError_t enqueue_eval(const Args<T> &A, int s, int control, hipStream_t st) {
    const int stats = control != 0;
    const unsigned gpart = (unsigned)num_partials(A);
    if (A.demag) {
        const size_t lx = lds_x(A), ly = lds_y(A);
        hipLaunchKernelGGL(k_fwd_x<T>, dim3(gpart), dim3(kThreads), lx, st, A, s);
        if (A.nz == 1) {
            hipLaunchKernelGGL((k_y<T, 0>), dim3(A.nkx), dim3(kThreads), ly, st, A);
        } else {
            hipLaunchKernelGGL((k_y<T, 1>), dim3(A.nkx * A.nz), dim3(kThreads), ly, st, A);
            const unsigned gz = (unsigned)(A.nkx * ((A.py + A.ty - 1) / A.ty));
            hipLaunchKernelGGL(k_z<T>, dim3(gz), dim3(kThreads), lds_z(A), st, A);
            hipLaunchKernelGGL((k_y<T, 2>), dim3(A.nkx * A.nz), dim3(kThreads), ly, st, A);
        }
        if (A.region) hipLaunchKernelGGL((k_inv_x<T, true>), dim3(gpart), dim3(kThreads), lx, st, A, s, stats, control);
        else hipLaunchKernelGGL((k_inv_x<T, false>), dim3(gpart), dim3(kThreads), lx, st, A, s, stats, control);
    } else {
        if (s >= 2) hipLaunchKernelGGL(k_stage<T>, dim3(gpart), dim3(kThreads), 0, st, A, s);
        if (A.region) hipLaunchKernelGGL((k_local<T, true>), dim3(gpart), dim3(kThreads), 0, st, A, s, stats, control);
        else hipLaunchKernelGGL((k_local<T, false>), dim3(gpart), dim3(kThreads), 0, st, A, s, stats, control);
    }
    return hipGetLastError();
}

template <class T>
hipError_t enqueue_control(const Args<T> &A, int observe, hipStream_t st) {
    hipLaunchKernelGGL(k_control<T>, dim3(1), dim3(kThreads), 0, st, A, (int)num_partials(A), observe);
    return hipGetLastError();
}

template const char *configure<float>(Args<float> &, size_t);
template const char *configure<double>(Args<double> &, size_t);
template long long num_partials<float>(const Args<float> &);
template long long num_partials<double>(const Args<double> &);
template hipError_t enqueue_eval<float>(const Args<float> &, int, int, hipStream_t);
template hipError_t enqueue_eval<double>(const Args<double> &, int, int, hipStream_t);
template hipError_t enqueue_control<float>(const Args<float> &, int, hipStream_t);
template hipError_t enqueue_control<double>(const Args<double> &, int, hipStream_t);

}  // namespace fm

namespace fm {

// One launch of a single kernel kind (for event timing): 0 x-forward or
// stage combine, 1 y pass, 2 z pass, 3 y inverse, 4 x-inverse/local pass,
// 5 controller.  Returns hipErrorInvalidValue for a kind the grid lacks.
template <class T>
hipError_t enqueue_kind(const Args<T> &A, int kind, int s, hipStream_t st) {
    const unsigned gpart = (unsigned)num_partials(A);
    const bool film = A.nz == 1;
    switch (kind) {
        case 0:
            if (A.demag) hipLaunchKernelGGL(k_fwd_x<T>, dim3(gpart), dim3(kThreads), lds_x(A), st, A, s);
            else hipLaunchKernelGGL(k_stage<T>, dim3(gpart), dim3(kThreads), 0, st, A, s);
            break;
        case 1:
            if (!A.demag) return hipErrorInvalidValue;
            if (film) hipLaunchKernelGGL((k_y<T, 0>), dim3(A.nkx), dim3(kThreads), lds_y(A), st, A);
            else hipLaunchKernelGGL((k_y<T, 1>), dim3(A.nkx * A.nz), dim3(kThreads), lds_y(A), st, A);
            break;
        case 2:
            if (!A.demag || film) return hipErrorInvalidValue;
            hipLaunchKernelGGL(k_z<T>, dim3((unsigned)(A.nkx * ((A.py + A.ty - 1) / A.ty))), dim3(kThreads), lds_z(A),
                               st, A);
            break;
        case 3:
            if (!A.demag || film) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_y<T, 2>), dim3(A.nkx * A.nz), dim3(kThreads), lds_y(A), st, A);
            break;
        case 4:
            if (A.demag) {
                if (A.region) hipLaunchKernelGGL((k_inv_x<T, true>), dim3(gpart), dim3(kThreads), lds_x(A), st, A, s, 1, 0);
                else hipLaunchKernelGGL((k_inv_x<T, false>), dim3(gpart), dim3(kThreads), lds_x(A), st, A, s, 1, 0);
            } else {
                if (A.region) hipLaunchKernelGGL((k_local<T, true>), dim3(gpart), dim3(kThreads), 0, st, A, s, 1, 0);
                else hipLaunchKernelGGL((k_local<T, false>), dim3(gpart), dim3(kThreads), 0, st, A, s, 1, 0);
            }
            break;
        case 5:
            hipLaunchKernelGGL(k_control<T>, dim3(1), dim3(kThreads), 0, st, A, (int)gpart, 0);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t enqueue_kind<float>(const Args<float> &, int, int, hipStream_t);
template hipError_t enqueue_kind<double>(const Args<double> &, int, int, hipStream_t);

}  // namespace fm
