// engine.hpp -- device state and kernel-argument layout of the MI355X FDM
// engine (shared by the HIP kernels and the C-ABI host layer).
//
// Data layout in HBM (SoA, cell index = (z*ny + y)*nx + x, as the reference
// exchange_fp64.cu:46-50):
//   m_all  : 2 slots x 3 x ncell  (ping-pong state; slot `parity` is current)
//   k_all  : 7 slots x 3 x ncell  (slots 0/1: k1 of the current state and
//            the last-stage RHS of the trial state; slots 2..6: k2..k6;
//            ABM3: slots 4..6 are the ring of f_n, f_{n-1}, f_{n-2})
//   ybuf   : 2 x 3 x ncell        (intermediate stage states, by stage parity)
//   spec   : 3 x nzs x nkx x ly complex (x-half-spectrum of Ms*m, y-lines
//            contiguous; nkx = nx+1, ly = ny for nz == 1 else 2ny, nzs = nz)
//   ntens  : real symmetric Newell spectrum, symmetry-reduced:
//            thin film: 4 comps (xx,yy,zz,xy) x nkx x (ny+1)
//            3-D:       nkx x (nz+1) x (ny+1) x 6 comps (xx,yy,zz,xy,xz,yz)
// The control block `Ctl` lives in device memory; every kernel reads dt,
// parity and the done flag from it, so a whole adaptive step (including
// rejections) runs without host round trips.
#pragma once
#include <cstdint>

namespace fm {

constexpr int kThreads = 256;       // work-group size of every kernel
constexpr int kMaxRowsWg = 4;       // x-rows owned by one x-pass work-group
constexpr int kMaxPasses = 24;      // FFT passes per length
constexpr int kMaxRadix = 64;       // largest prime factor of a padded length
constexpr int kStatSlots = 11;      // per-block partial reductions
constexpr int kMaxStages = 8;
constexpr int kMaxK = 5;            // k-vectors in one stage combination (DP45 rows 6, 7)
constexpr int kMaxErr = 6;          // terms of the embedded error estimate (DP45)
constexpr int kHist = 8;            // StageRow::kidx kHist + r names the ABM3 history value f_{n-r}
constexpr int kHistSlot = 4;        // first k_all slot of the ABM3 history ring
constexpr int kMaxLayers = 8;       // planes of a multilayer plan

// Per-block partial layout (doubles).
enum StatSlot {
    ST_ERR = 0,       // max error norm (max)
    ST_EEX = 1,       // exchange energy (sum)
    ST_EDEMAG = 2,    // demag energy (sum)
    ST_EEXT = 3,      // Zeeman energy (sum)
    ST_EANI = 4,      // anisotropy energy (sum)
    ST_HEFF2 = 5,     // max |H_eff|^2 (max)
    ST_HD2 = 6,       // max |H_d|^2 (max)
    ST_RHS2 = 7,      // max |dm/dt|^2 (max)
    ST_ECUB = 8,      // cubic anisotropy energy (sum)
    ST_EDMI = 9,      // DMI energy (sum)
    ST_EMEL = 10,     // magnetoelastic energy (sum)
};

struct FftDesc {
    int n;
    int lg;       // log2(n) when n is a power of two (radices 2/4/8 only), else -1
    int npass;
    int radix[kMaxPasses];
};

// Explicit Runge-Kutta tableau as stage rows.  Stage s (2..nstage) evaluates
// y_s = normalise(m0 + (dt*scale_s) * sum_q a_sq k_{j_sq}); stage `nstage`
// is the new state.  k_j for j == 1 is the RHS of m0, for j == nstage the RHS
// of the trial state.
struct StageRow {
    int nk;
    int kidx[kMaxStages];
    double a[kMaxStages];
    double scale;
};

struct Method {
    int nstage;         // index of the final stage (DP45: 7, RK4: 5, Heun: 3)
    int adaptive;       // 1 = embedded error estimate + step control
    double ctrl_exp;    // controller exponent 1/(order+1): DP45 1/5, RK23 1/3
    int nerr;
    int eidx[kMaxStages];
    double e[kMaxStages];
    // efuse = 1: the error terms other than the final stage's are exactly the
    // final row's k-vectors in row order (DP45), so the final combination
    // also forms their sum (Args::errv) and the last pass adds only k_final
    int efuse;
    // abm = 1: ABM3 predictor-corrector (stage 2's RHS is f*, stored in the
    // history slot it replaces); 2: its Heun start-up (the final RHS is also
    // pushed into the history).  0 for one-step methods.
    int abm;
    double node[kMaxStages + 1];   // stage s is evaluated at t + node[s]*dt
    StageRow row[kMaxStages + 1];  // row[s] for s = 2..nstage
};

// Time-dependent Zeeman waveforms f(t) (the field is H0 + f(t) H1).
enum Wave { WAVE_NONE = 0, WAVE_RAMP = 1, WAVE_SINC = 2, WAVE_SINE = 3 };

// Device-resident control block.
struct Ctl {
    double t;          // simulated time of the current state
    double dt;         // step size of the attempt in flight
    double dt_carry;   // runner dt carried between steps (loop mode)
    double t_end;      // loop mode: stop time
    double dt_used;    // last accepted dt
    double dt_next;    // controller proposal after the last accepted step
    double err;        // error norm of the last attempt
    double abm_dt;     // ABM3: dt of the last history push (0 = none)
    double stats[kStatSlots];  // reduced statistics of the last evaluation
    unsigned long long step;       // accepted steps since create
    unsigned long long step_end;   // loop mode: stop after this many steps
    unsigned long long rejected;   // rejected attempts since create
    unsigned long long attempts;   // attempts since create
    int parity;        // slot of the current state
    int accepted;      // last attempt accepted
    int done;          // loop mode finished: kernels return at once
    int loop;          // 1 = loop mode, 0 = single step driven by the host
    int abm_count;     // ABM3: history values since the last restart (0..3)
    int abm_head;      // ABM3: f_n is in k_all slot kHistSlot + abm_head
    int k1_redo;       // DP45 with a thermal field: k1 must be evaluated afresh
                       // (no FSAL stage: after create, upload or a rejection)
};

template <class T>
struct Args {
    // device copy of these arguments (kernels read AP[blockIdx.y]) and the
    // number of systems sharing a launch (batched sweeps; 1 otherwise)
    const Args *self_d;
    int nsys;
    // grid
    int nx, ny, nz;
    long long ncell;
    long long nrows;        // ny*nz x-lines
    int nkx, ly, nzs;       // spectrum extents
    int px, py, pz;         // padded FFT lengths (pz = 1 for thin film)
    int rows_per_wg;        // x-lines per work-group in the x passes
    int ty;                 // y-lines per work-group in the z pass
    int nseq_x;             // complex sequences per x work-group
    int xsb;                // LDS stride between those sequences (complex values; px for Stockham)
    double inv_p;           // 1 / (px*py*pz)
    // buffers
    T *m_all;
    T *k_all;
    T *ybuf;
    void *spec;             // complex<T>
    const T *ntens;
    const void *twx, *twy, *twz;   // complex<T> forward twiddles
    const void *tw4y;       // four-step twiddles W_py^{t1 k2} at [k2*R1 + t1] (thin-film fast y pass)
    const void *tw4y_f;     // 3-D forward y-lines: the same for the exchanged factors (R2, R1)
    int yfast;              // 0 = Stockham y pass, else py of the register four-step y pass
    const void *tw8;        // FP32 film at px or py = 1024: twiddle bases of the three-step FFTs ([6][128] complex), or null
    int yf8_allowed;        // 0 = keep the four-step y pass (FULLMAG_FDM_YPASS8=0)
    int yf8;                // 1 = the film y pass runs k_yf8
    const void *tw4x;       // four-step twiddles of the x passes (W_px^{t1 k2} at [k2*R1 + t1])
    int xfast;              // 0 = Stockham x passes, else px of the register four-step x passes
    int xpf;                // 1 = film joined x pass with its inputs prefetched into LDS (k_xxp)
    int xpf_off;            // byte offset of the prefetch area in the x passes' dynamic LDS
    int ylfast;             // 3-D: 0 = Stockham y passes, else py of the register four-step y-line passes
    int yl_threads;         // 3-D four-step y-line passes: threads per work-group (lines = yl_threads / max(R1,R2))
    int zf_ub;              // 3-D four-step z pass: consecutive ky per work-group
    const void *tw4z;       // four-step twiddles of the 3-D z pass (W_pz^{t1 k2} at [k2*R1 + t1])
    int zfast;              // 3-D: 0 = Stockham z pass, else pz of the register four-step z pass
    const uint8_t *mask;    // NULL = all active
    const uint8_t *nmask;   // with a mask: bit 0 = the cell, bit 1+q = its neighbour q
                            // (-x,+x,-y,+y,-z,+z as the stencil takes it) active
    const uint32_t *region; // NULL = uniform exchange
    const T *lut;           // 256 x 256 A_ij when region != NULL
    Ctl *ctl;
    double *partials;       // [nblocks][kStatSlots]
    // observables (written only when non-NULL)
    T *obs_hex, *obs_hd, *obs_heff;
    // adaptive methods with Method::efuse: sum over the final row of
    // e_q k_q (FP64, unscaled), written by the final stage combination
    double *errv;
    // trajectory record for the adjoint sweep (written only when non-NULL):
    // a ring of rec_seg step slots; for the step in flight, slot i =
    // (*rec_step mod rec_seg), stage state y_s at rec_y[i*rec_sv +
    // (s-1)*3N], effective field at y_s at rec_h[i*rec_sv + (s-1)*3N],
    // pre-normalisation norm |v_s| at rec_vn[i*rec_ss + (s-2)*N]; the final
    // stage's state and field (the next step's stage 1) start slot
    // (i + 1) mod rec_seg; the controller advances *rec_step
    T *rec_y, *rec_h, *rec_vn;
    long long *rec_step;
    long long rec_seg, rec_sv, rec_ss;
    // physics
    T ms;
    T ex_pre;               // 2A/(mu0 Ms)
    T lut_pre;              // 2/(mu0 Ms)
    T idx2, idy2, idz2;     // 1/dx^2 ...
    T hext[3];
    int has_ext;
    int has_uni;
    T uni_c1, uni_c2;       // 2Ku1/(mu0 Ms), 4Ku2/(mu0 Ms)
    T ku1_u, ku2_u;         // uniform Ku1, Ku2 (J/m^3)
    T axis[3];
    T gamma_bar, alpha;
    int precession;
    int demag;
    double e_pre;           // mu0*Ms*V (energy prefactor)
    // per-cell maps (NULL = the uniform value above)
    const T *ms_cell;       // Ms per cell
    const T *ms_pack;       // demag source weight per cell (Ms, or phi Ms with a boundary tier)
    const T *ku1_cell, *ku2_cell;
    const T *axis_cell;     // 3 x ncell SoA, unit vectors
    T two_a_mu0;            // 2A/mu0 (exchange prefactor numerator)
    T two_mu0inv;           // 2/mu0
    double mu0_vol;         // mu0*V
    // time-dependent Zeeman term f(t) H1 on the cells of wave_mask
    int wave_kind;
    double wave_amp[3], wave_t0, wave_t1, wave_freq, wave_phase;
    const uint8_t *wave_mask;
    // further terms (the kernels' XT instantiations evaluate them; xt = 0
    // when all are off).  Prefactors exclude 1/Ms, applied per cell.
    int xt;
    int has_cub;            // cubic anisotropy in the frame cub_ax = (c1, c2, c3)
    T cub_ax[9];
    T kc1_u, kc2_u, kc3_u;
    const T *kc1_cell, *kc2_cell, *kc3_cell;
    T dmi_i, dmi_b;         // interfacial / bulk D (J/m^2), 0 = off
    T i2x, i2y, i2z;        // 1/(2dx) ... (DMI central differences)
    T i1x, i1y, i1z;        // 1/dx ... (Zhang-Li upwind differences)
    int has_mel;            // magnetoelastic, uniform tensor strain e11 e22 e33 e23 e13 e12
    T mel_b1, mel_b2, mel_e[6];
    int has_zl;             // Zhang-Li: u = zl_u / Ms, non-adiabaticity zl_beta
    T zl_u[3], zl_beta;
    int has_slon;           // Slonczewski: pf/Ms * g(m.p) (m x (m x p) + eps m x p)
    T sl_p[3], sl_l2, sl_deg, sl_eps, sl_pf;
    int has_sot;            // spin-orbit: pf/Ms (-xi_dl m x (m x s) + xi_fl m x s)
    T sot_s[3], sot_pf, sot_dl, sot_fl;
    const T *oe_field;      // Oersted field at the full current, 3 x ncell SoA (NULL = off)
    int oe_kind;            // envelope: 0 constant, 1 sin(2 pi f t + phase) + offset, 2 pulse [t_on, t_off)
    double oe_freq, oe_phase, oe_offset, oe_t_on, oe_t_off;
    double cell_vol;        // dx dy dz
    int has_th;             // thermal field: sigma = sqrt(th_pre / (Ms dt)), dt = th_dt or the trial dt;
    double th_pre;          // 2 alpha kB T / (gamma (1 + alpha^2) mu0 V)
    double th_dt;           // 0 = the dt of the step attempt in flight
    unsigned long long th_seed;  // counter = accepted steps (Ctl::step)
    // boundary-correction tier: exchange stencil weights per cell in the
    // neighbour order -x, +x, -y, +y, -z, +z (6 x ncell, the 1/max(phi,
    // floor) normalisation included; NULL = standard stencil), volume
    // fraction (energy weights; NULL = 1), sparse demag correction per
    // target cell: sources dc_src[dc_ptr[c] .. dc_ptr[c+1]) with symmetric
    // tensors dc_n (6 per source, xx xy xz yy yz zz)
    const T *bc_w;
    const T *bc_phi;
    // periodic boundaries: bit a set = axis a wraps (neighbours of the
    // further-term instantiations); gtens = complex Newell spectrum of the
    // periodic kernel (thin film [c][kx][ky], 3-D [kx][kz][ky][c]; NULL =
    // the reduced real spectrum ntens)
    int pbc;
    const T *gtens;
    // multilayer: the nz planes are single-cell layers at arbitrary heights,
    // coupled only by demag through the pair kernels lmix
    // [dst][src][c][kx][ky] (c = xx,yy,zz,xy real, xz,yz imaginary parts;
    // kx <= nx, ky <= ny; NULL without demag or on a uniform grid);
    // layered = 1: no exchange/DMI/torque neighbours along z
    const T *lmix;
    int layered;
    const int *dc_ptr;
    const int *dc_src;
    const T *dc_n;
    // step control
    double atol, rtol, dt_min, dt_max, headroom, growth, shrink;
    int method;             // index into the constant method table
    int stage_nk[kMaxStages + 1];  // k-vectors combined by stage s (0 for s <= 1)
    FftDesc fx, fy, fz;
};

// Reverse (adjoint) sweep of one fixed step: cotangents and parameter
// gradients (see the adjoint kernels for the formulas).
template <class T>
struct AdjArgs {
    const T *rec_y, *rec_h, *rec_vn;  // record of the step being reversed (bound per launch by adj_bind)
    const T *rec_y0, *rec_h0, *rec_vn0;  // record bases (layout as in Args)
    long long *step;                  // global index n of the step being reversed (device)
    long long seg, sv, ss;            // record slots = n mod seg, slot strides
    const double *tn;                 // t_n for n = 0..steps (device)
    long long first;                  // objective_first_step
    long long obj_from;               // forward: control-block step count from which J accumulates
    double *cot_m;                    // 3N: cotangent of the step's start state (FP64: it accumulates over steps)
    double *cot_k;                    // kMaxStages x 3N: cotangent of k_j at [j*3N]
    double *beta;                     // 3N: field cotangent of the stage being reversed
    double *grad_ms;                  // N: dJ/dMs_i, accumulated
    double *grad_amp;                 // [work-groups][3]: dJ/d(wave amplitude), accumulated
    double *obj_part;                 // [work-groups]: objective partial sums, accumulated
    const T *obj_w;                   // N objective weights
    int obj_comp;                     // objective component (0 x, 1 y, 2 z)
    double obj_scale;                 // 1 / (number of objective samples)
    double t_n;                       // time at the start of the step
    double dt;                        // step size
};

// Size of the fixed device staging array of the allocation-free preview
// entry point (one per GPU): FP64 AoS bins of up to 87381 preview cells.
constexpr unsigned long long kPreviewScratchBytes = 2ull << 20;

// Field grid and preview request of the preview binning kernel.
struct PreviewGeom {
    uint32_t fx, fy, fz;     // field grid
    uint32_t pnx, pny, pnz;  // preview grid
    uint32_t z0, zs;         // z origin and z cells per preview plane
};

// Method indices in the constant table.
// M_ABM3S is the start-up table of M_ABM3 (selected on the device while the
// history holds fewer than three values).
enum MethodId { M_HEUN = 0, M_DP45 = 1, M_RK4 = 2, M_RK23 = 3, M_ABM3 = 4, M_ABM3S = 5, M_COUNT = 6 };

}  // namespace fm
