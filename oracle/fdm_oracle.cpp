// fdm_oracle.cpp -- FP64 CPU restatement of fullmag's `fdm_cpu_reference` path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the MI355X engine in
// fullmag_amd/: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  The product library never links it.
//
// It follows the reference's CPU engine (Rust, not buildable in this image):
//   Newell tensor ........ crates/fullmag-fdm-demag/src/newell.rs:21-517
//   FFT workspace ........ crates/fullmag-engine/src/fdm_fft.rs:64-131,445-602
//   exchange ............. crates/fullmag-engine/src/fdm_fields.rs:445-526
//   demag ................ crates/fullmag-engine/src/fdm_fields.rs:528-609
//   local terms .......... crates/fullmag-engine/src/fdm_fields.rs:910-1047
//   effective field ...... crates/fullmag-engine/src/fdm_fields.rs:1153-1183
//   observables/energies . crates/fullmag-engine/src/fdm_fields.rs:1683-1813,2168-2273,349-355
//   LLG RHS .............. crates/fullmag-engine/src/fdm_fields.rs:2151-2162
//   Heun ................. crates/fullmag-engine/src/fdm_integrators.rs:84-150
//   RK4 .................. crates/fullmag-engine/src/fdm_integrators.rs:155-264
//   RK45 (Dormand-Prince)  crates/fullmag-engine/src/fdm_integrators.rs:429-680
//   RK23 (Bogacki-Shampine) crates/fullmag-engine/src/fdm_integrators.rs:290-410
//   ABM3 (Adams-Bashforth-Moulton, Heun start-up)
//                           crates/fullmag-engine/src/fdm_integrators.rs:683-820,
//                           history: crates/fullmag-engine/src/fdm_state.rs:204-262
//   error norm ........... crates/fullmag-engine/src/fdm_integrators.rs:828-870
//   neighbour clamp ...... crates/fullmag-engine/src/fdm_types.rs:126-132
//   region exchange table  native/backends/fdm/src/exchange_fp64.cu:80-114 (per-neighbour
//                          A_ij = lut[r_i * 256 + r_j] from the region ids); the table the
//                          runner builds: crates/fullmag-runner/src/native_fdm.rs:215-238
//   cubic anisotropy ..... crates/fullmag-engine/src/fdm_fields.rs:648-705 (Kc1, Kc2);
//                          Kc3 term native/backends/fdm/src/demag_fp64.cu:301-330
//   interfacial/bulk DMI . crates/fullmag-engine/src/fdm_fields.rs:707-816
//   magnetoelastic ....... crates/fullmag-engine/src/magnetoelastic.rs:52-75,155-176
//   Oersted cylinder ..... crates/fullmag-engine/src/fdm_fields.rs:1049-1150 (geometry);
//                          time envelopes native/backends/fdm/src/demag_fp64.cu:628-650
//   thermal field ........ crates/fullmag-engine/src/fdm_fields.rs:818-905 (counter-based
//                          SplitMix64 keys, Box-Muller; counter advanced per step call,
//                          fdm_problem.rs:225-273)
//   Zhang-Li / Slonczewski / SOT torques
//                          crates/fullmag-engine/src/fdm_fields.rs:1446-1680, added to
//                          every RHS as native/backends/fdm/src/llg_fp64.cu:58-205 does
//                          (the CPU integrators leave them out of the stage RHS)
//   runner time loop ..... crates/fullmag-runner/src/cpu_reference.rs:448-455,622-639
//   random unit vectors .. crates/fullmag-plan/src/util.rs:118-147
//   hole mask ............ crates/fullmag-plan/src/geometry.rs:159-231
//
// The FFT is its own mixed-radix Stockham (the reference uses rustfft 6.4.1,
// whose exact rounding is not reproduced); it is pinned against numpy.fft and
// against a brute-force real-space convolution in tests/.
//
// Build: see oracle/Makefile (g++ -O3 -fopenmp -shared).

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

using cplx = std::complex<double>;
constexpr double PI = 3.14159265358979323846;
constexpr double MU0 = 4.0 * PI * 1e-7;
constexpr int ASYMPTOTIC_DISTANCE = 40;

// ---------------------------------------------------------------- Newell ---
// f and g base functions, newell.rs:21-122 (log1p form).
double newell_f(double x, double y, double z) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    if (r2 < 1e-300) return 0.0;
    const double r = std::sqrt(r2);
    double res = (2.0 * x2 - y2 - z2) * r / 6.0;
    const double rxz2 = x2 + z2;
    if (rxz2 > 1e-300) {
        const double arg = 2.0 * y * (y + r) / rxz2;
        if (arg > -1.0) res += y * (z2 - x2) / 4.0 * std::log1p(arg);
    }
    const double rxy2 = x2 + y2;
    if (rxy2 > 1e-300) {
        const double arg = 2.0 * z * (z + r) / rxy2;
        if (arg > -1.0) res += z * (y2 - x2) / 4.0 * std::log1p(arg);
    }
    if (std::fabs(x) > 1e-300) res -= x * y * z * std::atan(y * z / (x * r));
    return res;
}

double newell_g(double x, double y, double z) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    if (r2 < 1e-300) return 0.0;
    const double r = std::sqrt(r2);
    double res = -x * y * r / 3.0;
    const double rxy2 = x2 + y2;
    if (rxy2 > 1e-300) {
        const double arg = 2.0 * z * (z + r) / rxy2;
        if (arg > -1.0) res += x * y * z * std::log1p(arg) / 2.0;
    }
    const double ryz2 = y2 + z2;
    if (ryz2 > 1e-300) {
        const double arg = 2.0 * x * (x + r) / ryz2;
        if (arg > -1.0) res += y * (3.0 * z2 - y2) * std::log1p(arg) / 12.0;
    }
    const double rxz2 = x2 + z2;
    if (rxz2 > 1e-300) {
        const double arg = 2.0 * y * (y + r) / rxz2;
        if (arg > -1.0) res += x * (3.0 * z2 - x2) * std::log1p(arg) / 12.0;
    }
    if (std::fabs(z) > 1e-300) res -= z2 * z / 6.0 * std::atan(x * y / (z * r));
    if (std::fabs(y) > 1e-300) res -= y2 * z / 2.0 * std::atan(x * z / (y * r));
    if (std::fabs(x) > 1e-300) res -= x2 * z / 2.0 * std::atan(y * z / (x * r));
    return res;
}

// Kahan-Neumaier compensated sum, newell.rs:133-147.
double neumaier(const double *t, int n) {
    double s = 0.0, c = 0.0;
    for (int i = 0; i < n; ++i) {
        const double v = t[i];
        const double u = s + v;
        if (std::fabs(s) >= std::fabs(v)) c += (s - u) + v;
        else c += (v - u) + s;
        s = u;
    }
    return s + c;
}

// 27-point second-difference stencil of f (or g) at displacement (i,j,k),
// newell.rs:150-207.  fv holds values at signed indices -1..dist, offset +1.
double stencil27(int i, int j, int k, const std::vector<double> &fv, int sx, int sy, double hvol) {
    auto F = [&](int a, int b, int c) { return fv[(size_t)(c + 1) * sy * sx + (size_t)(b + 1) * sx + (a + 1)]; };
    const double t[27] = {
        8.0 * F(i, j, k),
        -4.0 * F(i + 1, j, k), -4.0 * F(i - 1, j, k),
        -4.0 * F(i, j + 1, k), -4.0 * F(i, j - 1, k),
        -4.0 * F(i, j, k + 1), -4.0 * F(i, j, k - 1),
        2.0 * F(i - 1, j - 1, k), 2.0 * F(i - 1, j + 1, k), 2.0 * F(i + 1, j - 1, k), 2.0 * F(i + 1, j + 1, k),
        2.0 * F(i - 1, j, k - 1), 2.0 * F(i - 1, j, k + 1), 2.0 * F(i + 1, j, k - 1), 2.0 * F(i + 1, j, k + 1),
        2.0 * F(i, j - 1, k - 1), 2.0 * F(i, j - 1, k + 1), 2.0 * F(i, j + 1, k - 1), 2.0 * F(i, j + 1, k + 1),
        -F(i - 1, j - 1, k - 1), -F(i - 1, j - 1, k + 1), -F(i - 1, j + 1, k - 1), -F(i + 1, j - 1, k - 1),
        -F(i - 1, j + 1, k + 1), -F(i + 1, j - 1, k + 1), -F(i + 1, j + 1, k - 1), -F(i + 1, j + 1, k + 1),
    };
    return neumaier(t, 27) / (4.0 * PI * hvol);
}

// Point-dipole far field, newell.rs:212-229.
double asym_diag(double x, double y, double z, double vol) {
    const double r2 = x * x + y * y + z * z;
    const double r = std::sqrt(r2);
    const double r3 = r2 * r;
    return (1.0 / r3 - 3.0 * x * x / (r3 * r2)) / (4.0 * PI) * vol;
}
double asym_off(double x, double y, double z, double vol) {
    const double r2 = x * x + y * y + z * z;
    const double r = std::sqrt(r2);
    const double r5 = r2 * r2 * r;
    return -3.0 * x * y / (4.0 * PI * r5) * vol;
}

// Real-space Newell tensor on the (2nx,2ny,2nz) padded grid with octant
// parity, newell.rs:254-517.  out: 6 arrays [xx,yy,zz,xy,xz,yz] of P each.
void newell_kernels(int nx, int ny, int nz, double dx, double dy, double dz, double *out) {
    const int px = 2 * nx, py = 2 * ny, pz = 2 * nz;
    const size_t P = (size_t)px * py * pz;
    const int fsx = std::min(nx, ASYMPTOTIC_DISTANCE) + 2;
    const int fsy = std::min(ny, ASYMPTOTIC_DISTANCE) + 2;
    const int fsz = std::min(nz, ASYMPTOTIC_DISTANCE) + 2;
    const size_t flen = (size_t)fsx * fsy * fsz;
    std::vector<double> fxx(flen), fyy(flen), fzz(flen), gxy(flen), gxz(flen), gyz(flen);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < fsz; ++k)
        for (int j = 0; j < fsy; ++j)
            for (int i = 0; i < fsx; ++i) {
                const double x = (i - 1) * dx, y = (j - 1) * dy, z = (k - 1) * dz;
                const size_t id = (size_t)k * fsy * fsx + (size_t)j * fsx + i;
                fxx[id] = newell_f(x, y, z);
                fyy[id] = newell_f(y, x, z);
                fzz[id] = newell_f(z, y, x);
                gxy[id] = newell_g(x, y, z);
                gxz[id] = newell_g(x, z, y);
                gyz[id] = newell_g(y, z, x);
            }
    std::memset(out, 0, sizeof(double) * 6 * P);
    double *nxx = out, *nyy = out + P, *nzz = out + 2 * P, *nxy = out + 3 * P, *nxz = out + 4 * P, *nyz = out + 5 * P;
    const double hvol = dx * dy * dz;
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double vxx, vyy, vzz, vxy, vxz, vyz;
                const long d2 = (long)i * i + (long)j * j + (long)k * k;
                const bool asym = i >= ASYMPTOTIC_DISTANCE || j >= ASYMPTOTIC_DISTANCE || k >= ASYMPTOTIC_DISTANCE ||
                                  d2 >= (long)ASYMPTOTIC_DISTANCE * ASYMPTOTIC_DISTANCE;
                if (asym) {
                    const double x = i * dx, y = j * dy, z = k * dz;
                    vxx = asym_diag(x, y, z, hvol);
                    vyy = asym_diag(y, x, z, hvol);
                    vzz = asym_diag(z, y, x, hvol);
                    vxy = asym_off(x, y, z, hvol);
                    vxz = asym_off(x, z, y, hvol);
                    vyz = asym_off(y, z, x, hvol);
                } else {
                    vxx = stencil27(i, j, k, fxx, fsx, fsy, hvol);
                    vyy = stencil27(i, j, k, fyy, fsx, fsy, hvol);
                    vzz = stencil27(i, j, k, fzz, fsx, fsy, hvol);
                    vxy = stencil27(i, j, k, gxy, fsx, fsy, hvol);
                    vxz = stencil27(i, j, k, gxz, fsx, fsy, hvol);
                    vyz = stencil27(i, j, k, gyz, fsx, fsy, hvol);
                }
                const int xs_n = i == 0 ? 1 : 2, ys_n = j == 0 ? 1 : 2, zs_n = k == 0 ? 1 : 2;
                const int xi[2] = {i, px - i}, yi[2] = {j, py - j}, zi[2] = {k, pz - k};
                const double sg[2] = {1.0, -1.0};
                for (int a = 0; a < xs_n; ++a)
                    for (int b = 0; b < ys_n; ++b)
                        for (int c = 0; c < zs_n; ++c) {
                            const size_t p = (size_t)zi[c] * py * px + (size_t)yi[b] * px + xi[a];
                            nxx[p] = vxx;
                            nyy[p] = vyy;
                            nzz[p] = vzz;
                            nxy[p] = vxy * sg[a] * sg[b];
                            nxz[p] = vxz * sg[a] * sg[c];
                            nyz[p] = vyz * sg[b] * sg[c];
                        }
            }
}

// ------------------------------------------------------------------- FFT ---
// Mixed-radix Stockham autosort FFT, unnormalised, sign -1 forward, +1 inverse.
struct FftPlan {
    int n = 0;
    std::vector<int> radix;
    std::vector<cplx> wf, wi;  // exp(-2 pi i k / n), exp(+2 pi i k / n)
    explicit FftPlan(int n_ = 1) : n(n_) {
        int m = n;
        while (m % 4 == 0) { radix.push_back(4); m /= 4; }
        while (m % 2 == 0) { radix.push_back(2); m /= 2; }
        for (int p = 3; m > 1; p += 2)
            while (m % p == 0) { radix.push_back(p); m /= p; }
        wf.resize(n);
        wi.resize(n);
        for (int k = 0; k < n; ++k) {
            const double a = 2.0 * PI * (double)k / (double)n;
            wf[k] = cplx(std::cos(a), -std::sin(a));
            wi[k] = std::conj(wf[k]);
        }
    }
    // data: n contiguous elements with element stride `stride`; scratch: 2n.
    void run(cplx *data, size_t stride, bool inverse, cplx *scratch) const {
        if (n == 1) return;
        cplx *a = scratch, *b = scratch + n;
        for (int i = 0; i < n; ++i) a[i] = data[(size_t)i * stride];
        const std::vector<cplx> &w = inverse ? wi : wf;
        int ns = 1;
        cplx v[64], o[64];
        std::vector<cplx> vbig, obig;
        for (int R : radix) {
            cplx *vv = v, *oo = o;
            if (R > 64) { vbig.resize(R); obig.resize(R); vv = vbig.data(); oo = obig.data(); }
            const int nr = n / R;
            const int tstep = n / (ns * R);
            for (int j = 0; j < nr; ++j) {
                const int k = j % ns;
                for (int r = 0; r < R; ++r) vv[r] = a[j + r * nr];
                if (ns > 1)
                    for (int r = 1; r < R; ++r) vv[r] *= w[(size_t)k * r * tstep % n];
                if (R == 2) {
                    oo[0] = vv[0] + vv[1];
                    oo[1] = vv[0] - vv[1];
                } else if (R == 4) {
                    const cplx s0 = vv[0] + vv[2], d0 = vv[0] - vv[2];
                    const cplx s1 = vv[1] + vv[3], d1 = vv[1] - vv[3];
                    // forward: -i * d1 ; inverse: +i * d1
                    const cplx jd1 = inverse ? cplx(-d1.imag(), d1.real()) : cplx(d1.imag(), -d1.real());
                    oo[0] = s0 + s1;
                    oo[2] = s0 - s1;
                    oo[1] = d0 + jd1;
                    oo[3] = d0 - jd1;
                } else {
                    const int step = n / R;
                    for (int s = 0; s < R; ++s) {
                        cplx acc = 0.0;
                        for (int q = 0; q < R; ++q) acc += vv[q] * w[(size_t)((q * s) % R) * step];
                        oo[s] = acc;
                    }
                }
                const int base = (j / ns) * ns * R + k;
                for (int r = 0; r < R; ++r) b[base + r * ns] = oo[r];
            }
            std::swap(a, b);
            ns *= R;
        }
        for (int i = 0; i < n; ++i) data[(size_t)i * stride] = a[i];
    }
};

// In-place 3-D FFT on a (px,py,pz) complex array, x then y then z
// (fdm_fft.rs:445-598 order).
void fft3(cplx *buf, int px, int py, int pz, const FftPlan &fx, const FftPlan &fy, const FftPlan &fz, bool inverse) {
    const int maxn = std::max(px, std::max(py, pz));
#pragma omp parallel
    {
        std::vector<cplx> scr(2 * (size_t)maxn);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)py * pz; ++r) fx.run(buf + (size_t)r * px, 1, inverse, scr.data());
        if (py > 1) {
#pragma omp for schedule(static)
            for (long r = 0; r < (long)px * pz; ++r) {
                const long z = r / px, x = r % px;
                fy.run(buf + (size_t)z * px * py + x, px, inverse, scr.data());
            }
        }
        if (pz > 1) {
#pragma omp for schedule(static)
            for (long r = 0; r < (long)px * py; ++r) fz.run(buf + r, (size_t)px * py, inverse, scr.data());
        }
    }
}

// ---------------------------------------------------------------- vector ---
struct V3 {
    double c[3];
};
inline V3 v_add(V3 a, V3 b) { return {{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
inline V3 v_scale(V3 a, double s) { return {{a.c[0] * s, a.c[1] * s, a.c[2] * s}}; }
inline double v_dot(V3 a, V3 b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2]; }
inline double v_norm(V3 a) { return std::sqrt(a.c[0] * a.c[0] + a.c[1] * a.c[1] + a.c[2] * a.c[2]); }
inline V3 v_cross(V3 a, V3 b) {
    return {{a.c[1] * b.c[2] - a.c[2] * b.c[1], a.c[2] * b.c[0] - a.c[0] * b.c[2], a.c[0] * b.c[1] - a.c[1] * b.c[0]}};
}
// vector.rs:11-19: normalise(0) = 0.
inline V3 v_normalized(V3 a) {
    const double n = v_norm(a);
    if (n <= 0.0) return {{0.0, 0.0, 0.0}};
    return v_scale(a, 1.0 / n);
}

}  // namespace

// ------------------------------------------------------------ public API ---
extern "C" {

typedef struct {
    uint32_t nx, ny, nz;
    double dx, dy, dz;
    double ms, a_ex, alpha, gamma;
    int precession;      // 0 = pure-damping RHS
    int exchange, demag;
    int has_ext;
    double h_ext[3];     // A/m
    int has_uni;
    double ku1, ku2, axis[3];
    const uint8_t *mask;  // NULL = all active
    double atol, rtol, dt_min, dt_max, headroom, growth, shrink;
    // Extensions without a reference counterpart (engine's plan extension):
    // per-cell Ms, Ku1, Ku2, anisotropy axis (NULL = uniform value above) and
    // a time-dependent Zeeman term H1 f(t) on the cells of wave_mask
    // (NULL = all active cells); f: 0 none, 1 ramp clamp((t-t0)/(t1-t0),0,1),
    // 2 sinc sin(x)/x with x = 2 pi freq (t-t0), 3 sin(2 pi freq t + phase).
    const double *ms_cell, *ku1_cell, *ku2_cell, *axis_cell;
    int wave_kind;
    double wave_amp[3];
    double wave_t0, wave_t1, wave_freq, wave_phase;
    const uint8_t *wave_mask;
    // cubic anisotropy: energy Kc1 S + Kc2 P + Kc3 S^2 in the frame c1, c2,
    // c3 = c1 x c2 (S = sum_{i<j} m_i^2 m_j^2, P = m1^2 m2^2 m3^2); per-cell
    // constants override the uniform ones
    int has_cub;
    double kc1, kc2, kc3, cub_axis1[3], cub_axis2[3];
    const double *kc1_cell, *kc2_cell, *kc3_cell;
    // DMI constants (J/m^2), 0 = off
    double dmi_i, dmi_b;
    // magnetoelastic coupling with a uniform prescribed strain, Voigt order
    // [e11, e22, e33, 2e23, 2e13, 2e12]
    int has_mel;
    double mel_b1, mel_b2, mel_strain[6];
    // current density (A/m^2) of the Zhang-Li and Slonczewski torques,
    // spin polarisation degree, non-adiabaticity; Slonczewski polarisation
    // axis, lambda and epsilon' (layer thickness dz)
    double j[3], stt_degree, stt_beta, slon_p[3], slon_lambda, slon_eps;
    // spin-orbit torque
    int has_sot;
    double sot_je, sot_xi_dl, sot_xi_fl, sot_sigma[3], sot_thickness;
    // Oersted field of an infinite cylinder; envelope 0 constant,
    // 1 sin(2 pi f t + phase) + offset, 2 pulse [t_on, t_off)
    int has_oe;
    double oe_current, oe_radius, oe_center[3], oe_axis[3];
    int oe_kind;
    double oe_freq, oe_phase, oe_offset, oe_t_on, oe_t_off;
    // thermal field: temperature (K), noise dt (0: the dt of the step in
    // flight), seed of the counter-based generator
    double temperature, thermal_dt;
    uint64_t thermal_seed;
    // boundary-correction tier: 0 none, 1 face-link weighted exchange (T0),
    // 2 intersection-distance stencil (T1); per-cell volume fraction phi
    // (required for a tier), face link fractions and distances to the
    // boundary in the order +x, -x, +y, -y, +z, -z (NULL = absent), phi
    // floor and distance minimum (0 = 0.05 and 0.1 min(dx, dy, dz))
    int bc_tier;
    double bc_phi_floor, bc_delta_min;
    const double *bc_phi;
    const double *bc_face[6];
    const double *bc_delta[6];
    // sparse demag correction (with a tier): dc_count targets, dc_stencil
    // sources each (-1 = none), tensors [xx, xy, xz, yy, yz, zz] per pair
    uint32_t dc_count, dc_stencil;
    const int32_t *dc_target, *dc_source;
    const double *dc_tensor;
    // periodic boundaries (FdmBoundaryPolicy, fdm_types.rs:94-132): bit 0/1/2
    // = x/y/z periodic; demag image counts per axis (0 = 10)
    uint32_t periodic;
    uint32_t images[3];
    // multilayer (multilayer.rs): the nz planes are single-cell layers with
    // centres at layer_z[k], coupled only by demag (NULL = uniform grid)
    const double *layer_z;
    // exchange region table (exchange_fp64.cu:80-114): region id per cell
    // (< 256) and the 256 x 256 stiffness table A_ij (J/m); NULL = the
    // uniform stiffness a_ex
    const uint32_t *region;
    const double *exchange_lut;
} orc_desc;

typedef struct {
    double time, dt_used, dt_next;
    double e_ex, e_demag, e_ext, e_ani, e_total;
    double max_heff, max_hdemag, max_rhs;
    uint64_t rejected;
    double e_cub, e_dmi, e_mel;  // cubic anisotropy, DMI, magnetoelastic (in e_total)
} orc_report;

}  // extern "C"

namespace {

struct Problem {
    orc_desc d;
    size_t n;
    std::vector<uint8_t> mask;
    bool has_mask;
    std::vector<double> ms_c, ku1_c, ku2_c, axis_c, kc1_c, kc2_c, kc3_c;
    std::vector<uint8_t> wmask;
    std::vector<V3> oe_prof;  // Oersted field at the full current, per cell
    std::vector<double> phi;                  // volume fraction (tier > 0)
    std::vector<double> face[6], dlt[6];      // T0 face links, T1 distances
    std::vector<int32_t> dc_t, dc_s;          // sparse demag correction
    std::vector<double> dc_n;
    std::vector<double> lz;                   // multilayer plane heights
    std::vector<uint32_t> reg;                // exchange region ids (empty: uniform stiffness)
    std::vector<double> lut;                  // 256 x 256 region stiffness table
    std::vector<std::vector<cplx>> lk;        // pair kernel spectra [(dst*L + src)*6 + c]
    uint64_t th_step = 0;     // thermal counter: step calls so far
    double th_dt_cur = 0.0;   // dt of the step attempt in flight
    double t_eval = 0.0;  // time at which the Zeeman waveform is evaluated
    // demag workspace (fdm_fft.rs:64-131)
    int px = 0, py = 0, pz = 0;
    size_t P = 0;
    FftPlan fx, fy, fz;
    std::vector<cplx> kxx, kyy, kzz, kxy, kxz, kyz;
    std::vector<cplx> bmx, bmy, bmz, bhx, bhy, bhz;
    // integrator buffers
    std::vector<V3> m0, ms, heff, hs, rhs;
    std::vector<V3> k[7];
    bool fsal_valid = false;
    std::vector<V3> kfsal;
    // ABM3 history (fdm_state.rs:204-262): f_n, f_{n-1}, f_{n-2}, pushes
    // since the last restart (saturating at 3) and the dt of the last push
    std::vector<V3> abm_f[3];
    int abm_count = 0;
    double abm_dt = 0.0;

    bool active(size_t i) const { return !has_mask || mask[i] != 0; }
    // exchange stiffness between cells i and j: A_{r_i r_j} from the region
    // table (exchange_fp64.cu:93-104), else the uniform a_ex
    double stiff(size_t i, size_t j) const { return reg.empty() ? d.a_ex : lut[(size_t)reg[i] * 256 + reg[j]]; }
    bool per(int a) const { return (d.periodic >> a) & 1u; }
    bool layered() const { return !lz.empty(); }
    // neighbor_index, fdm_types.rs:119-132: wrap on periodic axes, clamp
    // otherwise
    long nbr(int a, long i, long delta) const {
        const long n = a == 0 ? d.nx : a == 1 ? d.ny : d.nz;
        const long j = i + delta;
        if (a == 2 && layered()) return i;  // separate layers: no z neighbours
        if (per(a)) return (j % n + n) % n;
        return j < 0 ? 0 : j > n - 1 ? n - 1 : j;
    }
    // magnetic volume fraction of a cell (1 without a tier)
    double phi_at(size_t i) const { return phi.empty() ? 1.0 : phi[i]; }
    double ms_at(size_t i) const { return ms_c.empty() ? d.ms : ms_c[i]; }
    double wave(double t) const {
        switch (d.wave_kind) {
            case 1: {
                const double u = (t - d.wave_t0) / (d.wave_t1 - d.wave_t0);
                return u < 0.0 ? 0.0 : u > 1.0 ? 1.0 : u;
            }
            case 2: {
                const double x = 2.0 * PI * d.wave_freq * (t - d.wave_t0);
                return x == 0.0 ? 1.0 : std::sin(x) / x;
            }
            case 3: return std::sin(2.0 * PI * d.wave_freq * t + d.wave_phase);
            default: return 0.0;
        }
    }
};

// Periodic demag kernel by truncated images, fdm_fft.rs:611-717: on a
// periodic axis the workspace has n points (no padding) and
// K(i) = sum_{m=-I..I} N_open(i + m n) for i in [0, n), read from the open
// kernel of the grid extended to n (1 + 2 I) cells (its padded array indexed
// with wrap-around); open axes keep 2n padded points and m = 0.
void periodic_kernels(const Problem &p, double *out) {
    const int n[3] = {(int)p.d.nx, (int)p.d.ny, (int)p.d.nz};
    int I[3], en[3], ep[3];
    for (int a = 0; a < 3; ++a) {
        I[a] = p.per(a) ? (p.d.images[a] ? (int)p.d.images[a] : 10) : 0;
        en[a] = n[a] * (1 + 2 * I[a]);
        ep[a] = 2 * en[a];
    }
    const size_t EP = (size_t)ep[0] * ep[1] * ep[2];
    std::vector<double> open(6 * EP);
    newell_kernels(en[0], en[1], en[2], p.d.dx, p.d.dy, p.d.dz, open.data());
    const int px = p.px, py = p.py, pz = p.pz;
    const size_t P = (size_t)px * py * pz;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < pz; ++k)
        for (int j = 0; j < py; ++j)
            for (int i = 0; i < px; ++i) {
                double sum[6] = {0, 0, 0, 0, 0, 0};
                for (int mz = -I[2]; mz <= I[2]; ++mz)
                    for (int my = -I[1]; my <= I[1]; ++my)
                        for (int mx = -I[0]; mx <= I[0]; ++mx) {
                            auto wrap = [](long g, long e) { return (size_t)(((g % e) + e) % e); };
                            const size_t ei = wrap(i + (long)mx * n[0], ep[0]);
                            const size_t ej = wrap(j + (long)my * n[1], ep[1]);
                            const size_t ek = wrap(k + (long)mz * n[2], ep[2]);
                            const size_t src = ei + (size_t)ep[0] * (ej + (size_t)ep[1] * ek);
                            for (int c = 0; c < 6; ++c) sum[c] += open[c * EP + src];
                        }
                const size_t dst = (size_t)i + (size_t)px * (j + (size_t)py * k);
                for (int c = 0; c < 6; ++c) out[c * P + dst] = sum[c];
            }
}

// Multilayer demag, crates/fullmag-engine/src/multilayer.rs:129-260: every
// layer (one plane of nx x ny cells) is padded to (2nx, 2ny, 2) and
// transformed; the destination's spectrum is the sum over sources of the
// pair kernel times the source's, negated and transformed back.  Pair
// kernels: the exact self kernel for a layer with itself, otherwise the
// shifted kernel of newell.rs:528-660 / shifted_kernel.rs:24-46 for
// s = z_destination - z_source: N(i dx, j dy, s) from the 27-point stencil
// of f/g at z in {s - dz, s, s + dz} (point dipole where i or j >= 40 or
// i^2 + j^2 >= 40^2), placed with x/y parity at z index 0.
void shifted_layer_kernel(int nx, int ny, double dx, double dy, double dz, double s, double *out) {
    const int px = 2 * nx, py = 2 * ny, pz = 2;
    const size_t P = (size_t)px * py * pz;
    const int fsx = std::min(nx, ASYMPTOTIC_DISTANCE) + 2, fsy = std::min(ny, ASYMPTOTIC_DISTANCE) + 2, fsz = 3;
    const size_t flen = (size_t)fsx * fsy * fsz;
    std::vector<double> fxx(flen), fyy(flen), fzz(flen), gxy(flen), gxz(flen), gyz(flen);
    for (int k = 0; k < fsz; ++k)
        for (int j = 0; j < fsy; ++j)
            for (int i = 0; i < fsx; ++i) {
                const double x = (i - 1) * dx, y = (j - 1) * dy, z = (k - 1) * dz + s;
                const size_t id = (size_t)k * fsy * fsx + (size_t)j * fsx + i;
                fxx[id] = newell_f(x, y, z);
                fyy[id] = newell_f(y, x, z);
                fzz[id] = newell_f(z, y, x);
                gxy[id] = newell_g(x, y, z);
                gxz[id] = newell_g(x, z, y);
                gyz[id] = newell_g(y, z, x);
            }
    std::memset(out, 0, sizeof(double) * 6 * P);
    const double hvol = dx * dy * dz;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double v[6];
            const long d2 = (long)i * i + (long)j * j;
            if (i >= ASYMPTOTIC_DISTANCE || j >= ASYMPTOTIC_DISTANCE ||
                d2 >= (long)ASYMPTOTIC_DISTANCE * ASYMPTOTIC_DISTANCE) {
                const double x = i * dx, y = j * dy;
                v[0] = asym_diag(x, y, s, hvol);
                v[1] = asym_diag(y, x, s, hvol);
                v[2] = asym_diag(s, y, x, hvol);
                v[3] = asym_off(x, y, s, hvol);
                v[4] = asym_off(x, s, y, hvol);
                v[5] = asym_off(y, s, x, hvol);
            } else {
                v[0] = stencil27(i, j, 0, fxx, fsx, fsy, hvol);
                v[1] = stencil27(i, j, 0, fyy, fsx, fsy, hvol);
                v[2] = stencil27(i, j, 0, fzz, fsx, fsy, hvol);
                v[3] = stencil27(i, j, 0, gxy, fsx, fsy, hvol);
                v[4] = stencil27(i, j, 0, gxz, fsx, fsy, hvol);
                v[5] = stencil27(i, j, 0, gyz, fsx, fsy, hvol);
            }
            const int xs_n = i == 0 ? 1 : 2, ys_n = j == 0 ? 1 : 2;
            const int xi[2] = {i, px - i}, yi[2] = {j, py - j};
            const double sg[2] = {1.0, -1.0};
            for (int a = 0; a < xs_n; ++a)
                for (int b = 0; b < ys_n; ++b) {
                    const size_t q = (size_t)yi[b] * px + xi[a];
                    out[q] = v[0];
                    out[P + q] = v[1];
                    out[2 * P + q] = v[2];
                    out[3 * P + q] = v[3] * sg[a] * sg[b];
                    out[4 * P + q] = v[4] * sg[a];
                    out[5 * P + q] = v[5] * sg[b];
                }
        }
}

void layer_kernels(Problem &p) {
    const int nx = p.d.nx, ny = p.d.ny, L = p.d.nz;
    const size_t P = (size_t)p.px * p.py * p.pz;
    std::vector<double> self(6 * P), k(6 * P);
    newell_kernels(nx, ny, 1, p.d.dx, p.d.dy, p.d.dz, self.data());
    p.lk.assign((size_t)L * L * 6, {});
    for (int dst = 0; dst < L; ++dst)
        for (int src = 0; src < L; ++src) {
            const double *kr = self.data();
            if (dst != src) {
                shifted_layer_kernel(nx, ny, p.d.dx, p.d.dy, p.d.dz, p.lz[dst] - p.lz[src], k.data());
                kr = k.data();
            }
            for (int c = 0; c < 6; ++c) {
                auto &v = p.lk[((size_t)dst * L + src) * 6 + c];
                v.resize(P);
                for (size_t i = 0; i < P; ++i) v[i] = cplx(kr[c * P + i], 0.0);
                fft3(v.data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, false);
            }
        }
}

void demag_layers_add(Problem &p, const V3 *m, V3 *h) {
    const int nx = p.d.nx, ny = p.d.ny, L = p.d.nz;
    const size_t P = (size_t)p.px * p.py * p.pz, A = (size_t)nx * ny;
    std::vector<std::vector<cplx>> M((size_t)3 * L, std::vector<cplx>(P));
    for (int l = 0; l < L; ++l) {
        for (int c = 0; c < 3; ++c) std::fill(M[3 * l + c].begin(), M[3 * l + c].end(), cplx(0.0));
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t s = (size_t)l * A + (size_t)y * nx + x, t = (size_t)y * p.px + x;
                if (!p.active(s)) continue;
                const double ms = p.ms_at(s);
                for (int c = 0; c < 3; ++c) M[3 * l + c][t] = m[s].c[c] * ms;
            }
        for (int c = 0; c < 3; ++c) fft3(M[3 * l + c].data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, false);
    }
    std::vector<cplx> hx(P), hy(P), hz(P);
    const double norm = 1.0 / (double)P;
    for (int dst = 0; dst < L; ++dst) {
        std::fill(hx.begin(), hx.end(), cplx(0.0));
        std::fill(hy.begin(), hy.end(), cplx(0.0));
        std::fill(hz.begin(), hz.end(), cplx(0.0));
        for (int src = 0; src < L; ++src) {
            const auto *K = &p.lk[((size_t)dst * L + src) * 6];
            const auto &mx = M[3 * src], &my = M[3 * src + 1], &mz = M[3 * src + 2];
            for (size_t i = 0; i < P; ++i) {
                hx[i] += K[0][i] * mx[i] + K[3][i] * my[i] + K[4][i] * mz[i];
                hy[i] += K[3][i] * mx[i] + K[1][i] * my[i] + K[5][i] * mz[i];
                hz[i] += K[4][i] * mx[i] + K[5][i] * my[i] + K[2][i] * mz[i];
            }
        }
        for (auto *b : {&hx, &hy, &hz}) {
            for (auto &v : *b) v = -v;
            fft3(b->data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, true);
        }
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t s = (size_t)dst * A + (size_t)y * nx + x, t = (size_t)y * p.px + x;
                if (!p.active(s)) continue;
                h[s].c[0] += hx[t].real() * norm;
                h[s].c[1] += hy[t].real() * norm;
                h[s].c[2] += hz[t].real() * norm;
            }
    }
}

// Sparse demag correction, native/backends/fdm/src/demag_boundary_fp64.cu:23-76:
// H(target) -= sum over its sources of dN . (phi Ms m)(source), dN symmetric
// [xx, xy, xz, yy, yz, zz]; negative indices are skipped; repeated targets
// add up; inactive targets get no field.
void demag_correction_add(const Problem &p, const V3 *m, V3 *h) {
    const size_t S = p.dc_t.empty() ? 0 : p.dc_s.size() / p.dc_t.size();
    for (size_t t = 0; t < p.dc_t.size(); ++t) {
        const int32_t ti = p.dc_t[t];
        if (ti < 0 || !p.active((size_t)ti)) continue;
        double c[3] = {0.0, 0.0, 0.0};
        for (size_t j = 0; j < S; ++j) {
            const int32_t si = p.dc_s[t * S + j];
            if (si < 0) continue;
            const double *N = &p.dc_n[6 * (t * S + j)];
            const double w = p.phi_at((size_t)si) * p.ms_at((size_t)si);
            const double M[3] = {w * m[si].c[0], w * m[si].c[1], w * m[si].c[2]};
            c[0] -= N[0] * M[0] + N[1] * M[1] + N[2] * M[2];
            c[1] -= N[1] * M[0] + N[3] * M[1] + N[4] * M[2];
            c[2] -= N[2] * M[0] + N[4] * M[1] + N[5] * M[2];
        }
        for (int k = 0; k < 3; ++k) h[ti].c[k] += c[k];
    }
}

void demag_add(Problem &p, const V3 *m, V3 *h) {
    if (p.layered()) {
        demag_layers_add(p, m, h);
        return;
    }
    const int nx = p.d.nx, ny = p.d.ny, nz = p.d.nz;
    std::fill(p.bmx.begin(), p.bmx.end(), cplx(0.0));
    std::fill(p.bmy.begin(), p.bmy.end(), cplx(0.0));
    std::fill(p.bmz.begin(), p.bmz.end(), cplx(0.0));
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t s = ((size_t)z * ny + y) * nx + x;
                const size_t t = ((size_t)z * p.py + y) * p.px + x;
                if (!p.active(s)) continue;
                // with a tier the source is phi Ms m (demag_fp64.cu:81-86)
                const double ms = p.ms_at(s) * p.phi_at(s);
                p.bmx[t] = m[s].c[0] * ms;
                p.bmy[t] = m[s].c[1] * ms;
                p.bmz[t] = m[s].c[2] * ms;
            }
    fft3(p.bmx.data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, false);
    fft3(p.bmy.data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, false);
    fft3(p.bmz.data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, false);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)p.P; ++i) {
        const cplx mx = p.bmx[i], my = p.bmy[i], mz = p.bmz[i];
        p.bhx[i] = -(p.kxx[i] * mx + p.kxy[i] * my + p.kxz[i] * mz);
        p.bhy[i] = -(p.kxy[i] * mx + p.kyy[i] * my + p.kyz[i] * mz);
        p.bhz[i] = -(p.kxz[i] * mx + p.kyz[i] * my + p.kzz[i] * mz);
    }
    fft3(p.bhx.data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, true);
    fft3(p.bhy.data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, true);
    fft3(p.bhz.data(), p.px, p.py, p.pz, p.fx, p.fy, p.fz, true);
    const double norm = 1.0 / (double)p.P;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) {
        if (!p.active(s)) continue;
        const long x = s % nx, y = (s / nx) % ny, z = s / ((long)nx * ny);
        const size_t t = ((size_t)z * p.py + y) * p.px + x;
        h[s].c[0] += p.bhx[t].real() * norm;
        h[s].c[1] += p.bhy[t].real() * norm;
        h[s].c[2] += p.bhz[t].real() * norm;
    }
    demag_correction_add(p, m, h);
}

// Boundary-corrected exchange, native/backends/fdm/src/exchange_t0_fp64.cu:52-165
// (T0) and exchange_t1_fp64.cu:62-198 (T1), with the stiffness applied once
// (the reference kernels multiply by A in the sum and again in the prefactor
// when there is no region table) and the cell's own Ms.  Cells with phi <= 0
// or m = 0 get no exchange field; the sum is normalised by max(phi, floor).
// T0: sum over faces with link fraction f > 0 inside the box of
// f (m_j - m_i) / d^2.  T1: the neighbour on the -a side (inside the box,
// nonzero m) has weight 2 / (d (d + delta_{+a})), the one on the +a side
// 2 / (d (d + delta_{-a})), with delta clamped to delta_min.  z faces only
// for nz > 1.
void exchange_tier_add(const Problem &p, const V3 *m, V3 *h, int tier) {
    const long nx = p.d.nx, ny = p.d.ny, nz = p.d.nz;
    const double d[3] = {p.d.dx, p.d.dy, p.d.dz};
    const double floor_ = p.d.bc_phi_floor > 0.0 ? p.d.bc_phi_floor : 0.05;
    const double dmin = p.d.bc_delta_min > 0.0 ? p.d.bc_delta_min : 0.1 * std::min({p.d.dx, p.d.dy, p.d.dz});
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) {
        const double phi0 = p.phi[s];
        if (phi0 <= 0.0) continue;
        const V3 c = m[s];
        if (c.c[0] == 0.0 && c.c[1] == 0.0 && c.c[2] == 0.0) continue;
        const long xyz[3] = {s % nx, (s / nx) % ny, s / (nx * ny)};
        const long ext[3] = {nx, ny, nz};
        const long stride[3] = {1, nx, nx * ny};
        double acc[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a) {
            if (a == 2 && nz == 1) continue;
            for (int side = 0; side < 2; ++side) {  // 0: +a neighbour, 1: -a neighbour
                const bool inside = side == 0 ? xyz[a] + 1 < ext[a] : xyz[a] > 0;
                if (!inside) continue;
                const long q = side == 0 ? s + stride[a] : s - stride[a];
                double w;
                if (tier == 1) {
                    const double f = p.face[2 * a + side][s];
                    if (!(f > 0.0)) continue;
                    w = f / (d[a] * d[a]);
                } else {
                    const V3 v = m[q];
                    if (v.c[0] == 0.0 && v.c[1] == 0.0 && v.c[2] == 0.0) continue;
                    const double opp = std::max(p.dlt[2 * a + (1 - side)][s], dmin);
                    w = 2.0 / (d[a] * (d[a] + opp));
                }
                w *= p.stiff(s, q);
                for (int k = 0; k < 3; ++k) acc[k] += w * (m[q].c[k] - c.c[k]);
            }
        }
        const double pre = 2.0 / (MU0 * p.ms_at(s) * std::max(phi0, floor_));
        for (int k = 0; k < 3; ++k) h[s].c[k] += pre * acc[k];
    }
}

// The stencil tier in effect: T0 needs all six face-link arrays, T1 all six
// distance arrays (exchange_fp64.cu:162-205); otherwise the standard
// exchange runs.
int exchange_tier(const Problem &p) {
    if (p.d.bc_tier == 1 && !p.face[0].empty()) return 1;
    if (p.d.bc_tier >= 2 && !p.dlt[0].empty()) return 2;
    return 0;
}

void exchange_add(const Problem &p, const V3 *m, V3 *h) {
    if (const int tier = exchange_tier(p)) {
        exchange_tier_add(p, m, h, tier);
        return;
    }
    const double dx2 = p.d.dx * p.d.dx, dy2 = p.d.dy * p.d.dy, dz2 = p.d.dz * p.d.dz;
    const long nx = p.d.nx, ny = p.d.ny;
    if (!p.reg.empty()) {
        // region table form (exchange_fp64.cu:80-114): H = 2/(mu0 Ms) sum_j
        // A_{r_i r_j} (m_j - m_i)/d^2 over the clamped (or wrapped)
        // neighbours; an inactive neighbour contributes nothing
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)p.n; ++s) {
            if (!p.active(s)) continue;
            const long x = s % nx, y = (s / nx) % ny, z = s / (nx * ny);
            const long nb[6] = {(z * ny + y) * nx + p.nbr(0, x, -1), (z * ny + y) * nx + p.nbr(0, x, 1),
                                (z * ny + p.nbr(1, y, -1)) * nx + x, (z * ny + p.nbr(1, y, 1)) * nx + x,
                                (p.nbr(2, z, -1) * ny + y) * nx + x, (p.nbr(2, z, 1) * ny + y) * nx + x};
            const double w[6] = {1.0 / dx2, 1.0 / dx2, 1.0 / dy2, 1.0 / dy2, 1.0 / dz2, 1.0 / dz2};
            double acc[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < 6; ++q) {
                if (!p.active(nb[q])) continue;
                const double a = p.stiff(s, nb[q]) * w[q];
                for (int k = 0; k < 3; ++k) acc[k] += a * (m[nb[q]].c[k] - m[s].c[k]);
            }
            const double pre = 2.0 / (MU0 * p.ms_at(s));
            for (int k = 0; k < 3; ++k) h[s].c[k] += pre * acc[k];
        }
        return;
    }
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) {
        if (!p.active(s)) continue;
        const double pre = 2.0 * p.d.a_ex / (MU0 * p.ms_at(s));
        const long x = s % nx, y = (s / nx) % ny, z = s / (nx * ny);
        const V3 c = m[s];
        auto sample = [&](long xi, long yi, long zi) {
            const size_t q = ((size_t)zi * ny + yi) * nx + xi;
            return p.active(q) ? m[q] : c;
        };
        // fdm_types.rs:119-132 clamp or periodic wrap
        const V3 xm = sample(p.nbr(0, x, -1), y, z), xp = sample(p.nbr(0, x, 1), y, z);
        const V3 ym = sample(x, p.nbr(1, y, -1), z), yp = sample(x, p.nbr(1, y, 1), z);
        const V3 zm = sample(x, y, p.nbr(2, z, -1)), zp = sample(x, y, p.nbr(2, z, 1));
        for (int k = 0; k < 3; ++k)
            h[s].c[k] += pre * ((xp.c[k] - 2.0 * c.c[k] + xm.c[k]) / dx2 + (yp.c[k] - 2.0 * c.c[k] + ym.c[k]) / dy2 +
                                (zp.c[k] - 2.0 * c.c[k] + zm.c[k]) / dz2);
    }
}

void ext_add(const Problem &p, V3 *h) {
    const double f = p.d.wave_kind ? p.wave(p.t_eval) : 0.0;
    if (!p.d.has_ext && !p.d.wave_kind) return;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s)
        if (p.active(s))
            for (int k = 0; k < 3; ++k) {
                if (p.d.has_ext) h[s].c[k] += p.d.h_ext[k];
                if (p.d.wave_kind && (p.wmask.empty() || p.wmask[s])) h[s].c[k] += f * p.d.wave_amp[k];
            }
}

void uni_add(const Problem &p, const V3 *m, V3 *h) {
    if (!p.d.has_uni) return;
    V3 u0 = {{p.d.axis[0], p.d.axis[1], p.d.axis[2]}};
    u0 = v_scale(u0, 1.0 / std::max(v_norm(u0), 1e-30));
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) {
        if (!p.active(s)) continue;
        const double ms = std::max(p.ms_at(s), 1e-30);
        V3 u = u0;
        if (!p.axis_c.empty()) {
            u = {{p.axis_c[3 * s], p.axis_c[3 * s + 1], p.axis_c[3 * s + 2]}};
            u = v_scale(u, 1.0 / std::max(v_norm(u), 1e-30));
        }
        const double k1 = p.ku1_c.empty() ? p.d.ku1 : p.ku1_c[s];
        const double k2 = p.ku2_c.empty() ? p.d.ku2 : p.ku2_c[s];
        const double mu = v_dot(m[s], u);
        const double coeff = 2.0 * k1 / (MU0 * ms) * mu + 4.0 * k2 / (MU0 * ms) * mu * mu * mu;
        for (int k = 0; k < 3; ++k) h[s].c[k] += u.c[k] * coeff;
    }
}

V3 unit_or_zero(const double *a) {
    V3 u = {{a[0], a[1], a[2]}};
    return v_scale(u, 1.0 / std::max(v_norm(u), 1e-30));
}

// Cubic anisotropy, fdm_fields.rs:648-705 (axes normalised, c3 = c1 x c2),
// plus the Kc3 S^2 term of demag_fp64.cu:301-330.
void cubic_add(const Problem &p, const V3 *m, V3 *h) {
    if (!p.d.has_cub) return;
    const V3 c1 = unit_or_zero(p.d.cub_axis1), c2 = unit_or_zero(p.d.cub_axis2), c3 = v_cross(c1, c2);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) {
        if (!p.active(s)) continue;
        const double ms = std::max(p.ms_at(s), 1e-30);
        const double k1 = p.kc1_c.empty() ? p.d.kc1 : p.kc1_c[s];
        const double k2 = p.kc2_c.empty() ? p.d.kc2 : p.kc2_c[s];
        const double k3 = p.kc3_c.empty() ? p.d.kc3 : p.kc3_c[s];
        const double m1 = v_dot(m[s], c1), m2 = v_dot(m[s], c2), m3 = v_dot(m[s], c3);
        const double a1 = m1 * m1, a2 = m2 * m2, a3 = m3 * m3;
        const double sig = a1 * a2 + a2 * a3 + a1 * a3;
        const double pf = 2.0 / (MU0 * ms), pf3 = 4.0 * k3 / (MU0 * ms);
        const double g1 = -pf * (k1 * m1 * (a2 + a3) + k2 * m1 * a2 * a3) - pf3 * sig * m1 * (a2 + a3);
        const double g2 = -pf * (k1 * m2 * (a1 + a3) + k2 * m2 * a1 * a3) - pf3 * sig * m2 * (a1 + a3);
        const double g3 = -pf * (k1 * m3 * (a1 + a2) + k2 * m3 * a1 * a2) - pf3 * sig * m3 * (a1 + a2);
        for (int k = 0; k < 3; ++k) h[s].c[k] += g1 * c1.c[k] + g2 * c2.c[k] + g3 * c3.c[k];
    }
}

// Interfacial and bulk DMI, fdm_fields.rs:707-816: central differences over
// box-clamped neighbours (fdm_types.rs:126-132), neighbour values as stored
// (zero on empty cells); H_i = 2D/(mu0 Ms) (dz mz... ), H_b = -2D/(mu0 Ms) curl m.
void dmi_add(const Problem &p, const V3 *m, V3 *h) {
    if (p.d.dmi_i == 0.0 && p.d.dmi_b == 0.0) return;
    const long nx = p.d.nx, ny = p.d.ny;
    const double i2x = 1.0 / (2.0 * p.d.dx), i2y = 1.0 / (2.0 * p.d.dy), i2z = 1.0 / (2.0 * p.d.dz);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) {
        if (!p.active(s)) continue;
        const double ms = std::max(p.ms_at(s), 1e-30);
        const long x = s % nx, y = (s / nx) % ny, z = s / (nx * ny);
        auto at = [&](long xi, long yi, long zi) { return m[((size_t)zi * ny + yi) * nx + xi]; };
        const V3 xm = at(p.nbr(0, x, -1), y, z), xp = at(p.nbr(0, x, 1), y, z);
        const V3 ym = at(x, p.nbr(1, y, -1), z), yp = at(x, p.nbr(1, y, 1), z);
        const V3 zm = at(x, y, p.nbr(2, z, -1)), zp = at(x, y, p.nbr(2, z, 1));
        if (p.d.dmi_i != 0.0) {
            const double pf = 2.0 * p.d.dmi_i / (MU0 * ms);
            h[s].c[0] += pf * (xp.c[2] - xm.c[2]) * i2x;
            h[s].c[1] += pf * (yp.c[2] - ym.c[2]) * i2y;
            h[s].c[2] += -pf * ((xp.c[0] - xm.c[0]) * i2x + (yp.c[1] - ym.c[1]) * i2y);
        }
        if (p.d.dmi_b != 0.0) {
            const double pf = -2.0 * p.d.dmi_b / (MU0 * ms);
            h[s].c[0] += pf * ((yp.c[2] - ym.c[2]) * i2y - (zp.c[1] - zm.c[1]) * i2z);
            h[s].c[1] += pf * ((zp.c[0] - zm.c[0]) * i2z - (xp.c[2] - xm.c[2]) * i2x);
            h[s].c[2] += pf * ((xp.c[1] - xm.c[1]) * i2x - (yp.c[0] - ym.c[0]) * i2y);
        }
    }
}

// Magnetoelastic field and energy density, magnetoelastic.rs:52-75,155-176.
V3 mel_field(const Problem &p, V3 m, double ms) {
    const double *e = p.d.mel_strain;
    const double e12 = 0.5 * e[5], e13 = 0.5 * e[4], e23 = 0.5 * e[3];
    const double b1 = p.d.mel_b1, b2 = p.d.mel_b2, inv = -1.0 / (MU0 * ms);
    return {{inv * (2.0 * b1 * m.c[0] * e[0] + 2.0 * b2 * (m.c[1] * e12 + m.c[2] * e13)),
             inv * (2.0 * b1 * m.c[1] * e[1] + 2.0 * b2 * (m.c[0] * e12 + m.c[2] * e23)),
             inv * (2.0 * b1 * m.c[2] * e[2] + 2.0 * b2 * (m.c[0] * e13 + m.c[1] * e23))}};
}
double mel_density(const Problem &p, V3 m) {
    const double *e = p.d.mel_strain;
    const double e12 = 0.5 * e[5], e13 = 0.5 * e[4], e23 = 0.5 * e[3];
    return p.d.mel_b1 * (m.c[0] * m.c[0] * e[0] + m.c[1] * m.c[1] * e[1] + m.c[2] * m.c[2] * e[2]) +
           2.0 * p.d.mel_b2 * (m.c[0] * m.c[1] * e12 + m.c[0] * m.c[2] * e13 + m.c[1] * m.c[2] * e23);
}
void mel_add(const Problem &p, const V3 *m, V3 *h) {
    if (!p.d.has_mel) return;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s)
        if (p.active(s)) h[s] = v_add(h[s], mel_field(p, m[s], std::max(p.ms_at(s), 1e-30)));
}

// Oersted profile at the full current, fdm_fields.rs:1049-1150 (cell
// centres, radial distance from the axis line, azimuthal axis x r_hat).
void oe_profile(Problem &p) {
    const long nx = p.d.nx, ny = p.d.ny;
    const V3 ax = unit_or_zero(p.d.oe_axis);
    const double I = p.d.oe_current, R = p.d.oe_radius;
    p.oe_prof.assign(p.n, V3{{0.0, 0.0, 0.0}});
    for (size_t s = 0; s < p.n; ++s) {
        if (!p.active(s)) continue;
        const long x = (long)s % nx, y = ((long)s / nx) % ny, z = (long)s / (nx * ny);
        const V3 r0 = {{(x + 0.5) * p.d.dx - p.d.oe_center[0], (y + 0.5) * p.d.dy - p.d.oe_center[1],
                        (z + 0.5) * p.d.dz - p.d.oe_center[2]}};
        const double proj = v_dot(r0, ax);
        const V3 r = v_add(r0, v_scale(ax, -proj));
        const double rp = v_norm(r);
        if (rp < 1e-30) continue;
        const V3 phi = v_cross(ax, v_scale(r, 1.0 / rp));
        const double hm = rp <= R ? I * rp / (2.0 * PI * R * R) : I / (2.0 * PI * rp);
        p.oe_prof[s] = v_scale(phi, hm);
    }
}
double oe_envelope(const Problem &p, double t) {
    switch (p.d.oe_kind) {
        case 1: return std::sin(2.0 * PI * p.d.oe_freq * t + p.d.oe_phase) + p.d.oe_offset;
        case 2: return (t >= p.d.oe_t_on && t < p.d.oe_t_off) ? 1.0 : 0.0;
        default: return 1.0;
    }
}
void oe_add(const Problem &p, V3 *h) {
    if (p.oe_prof.empty()) return;
    const double f = oe_envelope(p, p.t_eval);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) h[s] = v_add(h[s], v_scale(p.oe_prof[s], f));
}

// Zhang-Li (upwind differences), Slonczewski and spin-orbit torques,
// fdm_fields.rs:1446-1680, activation as native/backends/fdm/src/api.cpp:250-303.
void torque_add(const Problem &p, const V3 *m, V3 *k) {
    const double jm = std::sqrt(p.d.j[0] * p.d.j[0] + p.d.j[1] * p.d.j[1] + p.d.j[2] * p.d.j[2]);
    const bool zl = jm > 0.0 && p.d.stt_degree > 0.0;
    const double pp = v_dot(V3{{p.d.slon_p[0], p.d.slon_p[1], p.d.slon_p[2]}}, V3{{p.d.slon_p[0], p.d.slon_p[1], p.d.slon_p[2]}});
    const bool sl = pp > 0.0 && p.d.slon_lambda > 0.0 && jm > 0.0;
    const bool so = p.d.has_sot && p.d.sot_je != 0.0 && p.d.sot_thickness > 0.0;
    if (!zl && !sl && !so) return;
    constexpr double MU_B = 9.274009994e-24, E_CH = 1.60217662e-19, HBAR = 1.054571817e-34;
    const long nx = p.d.nx, ny = p.d.ny, nz = p.d.nz;
    const V3 sg = unit_or_zero(p.d.sot_sigma);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) {
        if (!p.active(s)) continue;
        const double ms = std::max(p.ms_at(s), 1e-30);
        const V3 mm = m[s];
        if (zl) {
            const double beta = p.d.stt_beta;
            const double b = p.d.stt_degree * MU_B / (E_CH * ms * (1.0 + beta * beta));
            const long x = s % nx, y = (s / nx) % ny, z = s / (nx * ny);
            const long st[3] = {1, nx, nx * ny}, pos[3] = {x, y, z}, len[3] = {nx, ny, nz};
            const double dd[3] = {p.d.dx, p.d.dy, p.d.dz};
            V3 du = {{0.0, 0.0, 0.0}};
            for (int a = 0; a < 3; ++a) {
                if (a == 2 && p.layered()) continue;
                const double u = b * p.d.j[a];
                // upwind difference; a box face contributes nothing unless
                // the axis is periodic (fdm_fields.rs:1290-1330)
                const long prev = s + (p.nbr(a, pos[a], -1) - pos[a]) * st[a];
                const long next = s + (p.nbr(a, pos[a], 1) - pos[a]) * st[a];
                if (u > 0.0 && (p.per(a) || pos[a] > 0)) du = v_add(du, v_scale(v_add(mm, v_scale(m[prev], -1.0)), u / dd[a]));
                else if (u < 0.0 && (p.per(a) || pos[a] + 1 < len[a]))
                    du = v_add(du, v_scale(v_add(m[next], v_scale(mm, -1.0)), u / dd[a]));
            }
            const V3 c = v_cross(mm, du), dc = v_cross(mm, c);
            k[s] = v_add(k[s], v_add(v_scale(dc, -1.0), v_scale(c, -beta)));
        }
        if (sl) {
            const V3 pv = {{p.d.slon_p[0], p.d.slon_p[1], p.d.slon_p[2]}};
            const double pre = jm * HBAR / (2.0 * E_CH * MU0 * ms * std::max(p.d.dz, 1e-30));
            const double l2 = p.d.slon_lambda * p.d.slon_lambda;
            const double deg = p.d.stt_degree > 0.0 ? p.d.stt_degree : 1.0;
            const double g = deg * l2 / ((l2 + 1.0) + (l2 - 1.0) * v_dot(mm, pv));
            const V3 mp = v_cross(mm, pv), mmp = v_cross(mm, mp);
            k[s] = v_add(k[s], v_scale(v_add(mmp, v_scale(mp, p.d.slon_eps)), pre * g));
        }
        if (so) {
            const double amp = std::fabs(p.d.sot_je) * HBAR / (2.0 * E_CH * MU0 * ms * p.d.sot_thickness);
            const V3 ms_ = v_cross(mm, sg), mms = v_cross(mm, ms_);
            k[s] = v_add(k[s], v_scale(v_add(v_scale(mms, -p.d.sot_xi_dl), v_scale(ms_, p.d.sot_xi_fl)), amp));
        }
    }
}

// Thermal field, fdm_fields.rs:818-905: per cell three Gaussians from four
// uniforms of the SplitMix64 finaliser applied to a key of (seed, counter,
// cell, stream); sigma = sqrt(2 alpha kB T / (gamma (1 + alpha^2) mu0 Ms V dt)).
uint64_t splitmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double counter_uniform(uint64_t seed, uint64_t step, uint64_t cell, uint64_t stream) {
    const uint64_t key = seed + step * 0x9E3779B97F4A7C15ull + cell * 0x517CC1B727220A95ull +
                         stream * 0x6C62272E07BB0142ull;
    return ((double)(splitmix64(key) >> 11) + 1.0) / ((double)(1ull << 53) + 1.0);
}
void thermal_add(const Problem &p, V3 *h) {
    const double dt = p.d.thermal_dt > 0.0 ? p.d.thermal_dt : p.th_dt_cur;
    if (!(p.d.temperature > 0.0) || !(dt > 0.0)) return;
    constexpr double KB = 1.380649e-23;
    const double a = p.d.alpha, g0 = p.d.gamma * (1.0 + a * a), vol = p.d.dx * p.d.dy * p.d.dz;
    const uint64_t seed = p.d.thermal_seed ? p.d.thermal_seed : 42;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) {
        if (!p.active(s)) continue;
        const double sig = std::sqrt(2.0 * a * KB * p.d.temperature / (g0 * MU0 * p.ms_at(s) * vol * dt));
        const uint64_t c = (uint64_t)s;
        const double u1 = std::max(counter_uniform(seed, p.th_step, c, 0), 1e-300);
        const double u2 = counter_uniform(seed, p.th_step, c, 1);
        const double u3 = std::max(counter_uniform(seed, p.th_step, c, 2), 1e-300);
        const double u4 = counter_uniform(seed, p.th_step, c, 3);
        const double r1 = std::sqrt(-2.0 * std::log(u1)), r2 = std::sqrt(-2.0 * std::log(u3));
        h[s].c[0] += sig * r1 * std::cos(2.0 * PI * u2);
        h[s].c[1] += sig * r1 * std::sin(2.0 * PI * u2);
        h[s].c[2] += sig * r2 * std::cos(2.0 * PI * u4);
    }
}

// fdm_fields.rs:1153-1183; the Oersted envelope and the torques of
// torque_add as the reference's GPU engine applies them.
void effective_field(Problem &p, const V3 *m, V3 *h) {
    std::fill(h, h + p.n, V3{{0.0, 0.0, 0.0}});
    if (p.d.exchange) exchange_add(p, m, h);
    if (p.d.demag) demag_add(p, m, h);
    mel_add(p, m, h);
    ext_add(p, h);
    uni_add(p, m, h);
    cubic_add(p, m, h);
    thermal_add(p, h);
    dmi_add(p, m, h);
    oe_add(p, h);
}

inline V3 llg(const Problem &p, V3 m, V3 h) {
    const double a = p.d.alpha;
    const double gb = p.d.gamma / (1.0 + a * a);
    const V3 pr = v_cross(m, h);
    const V3 dm = v_cross(m, pr);
    const V3 pt = p.d.precession ? pr : V3{{0.0, 0.0, 0.0}};
    return v_scale(v_add(pt, v_scale(dm, a)), -gb);
}

void rhs_eval(Problem &p, const V3 *m, V3 *k) {
    effective_field(p, m, p.heff.data());
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)p.n; ++s) k[s] = llg(p, m[s], p.heff[s]);
    torque_add(p, m, k);
}

double max_norm(const V3 *v, size_t n) {
    double r = 0.0;
#pragma omp parallel for reduction(max : r) schedule(static)
    for (long s = 0; s < (long)n; ++s) r = std::max(r, v_norm(v[s]));
    return r;
}

// With a tier every cell's energy is weighted by its volume fraction.
double energy_field_form(const Problem &p, const V3 *m, const V3 *h, double f) {
    const double vol = p.d.dx * p.d.dy * p.d.dz;
    double e = 0.0;
    for (size_t s = 0; s < p.n; ++s) e += f * MU0 * p.ms_at(s) * p.phi_at(s) * v_dot(m[s], h[s]) * vol;
    return e;
}

// compute_step_observables_zero_alloc, fdm_fields.rs:1683-1813.
void observables(Problem &p, const V3 *m, orc_report *r) {
    const size_t n = p.n;
    std::fill(p.heff.begin(), p.heff.end(), V3{{0.0, 0.0, 0.0}});
    r->e_ex = r->e_demag = r->e_ext = r->e_ani = 0.0;
    r->max_hdemag = 0.0;
    if (p.d.exchange) {
        std::fill(p.hs.begin(), p.hs.end(), V3{{0.0, 0.0, 0.0}});
        exchange_add(p, m, p.hs.data());
        r->e_ex = energy_field_form(p, m, p.hs.data(), -0.5);
        for (size_t s = 0; s < n; ++s) p.heff[s] = v_add(p.heff[s], p.hs[s]);
    }
    if (p.d.demag) {
        std::fill(p.hs.begin(), p.hs.end(), V3{{0.0, 0.0, 0.0}});
        demag_add(p, m, p.hs.data());
        r->e_demag = energy_field_form(p, m, p.hs.data(), -0.5);
        r->max_hdemag = max_norm(p.hs.data(), n);
        for (size_t s = 0; s < n; ++s) p.heff[s] = v_add(p.heff[s], p.hs[s]);
    }
    if (p.d.has_ext || p.d.wave_kind) {
        std::fill(p.hs.begin(), p.hs.end(), V3{{0.0, 0.0, 0.0}});
        ext_add(p, p.hs.data());
        r->e_ext = energy_field_form(p, m, p.hs.data(), -1.0);
        for (size_t s = 0; s < n; ++s) p.heff[s] = v_add(p.heff[s], p.hs[s]);
    }
    // uniaxial, cubic and DMI energies in field form (-1/2 mu0 Ms m.H V),
    // magnetoelastic from its density; all of them in the total
    std::fill(p.hs.begin(), p.hs.end(), V3{{0.0, 0.0, 0.0}});
    uni_add(p, m, p.hs.data());
    r->e_ani = energy_field_form(p, m, p.hs.data(), -0.5);
    for (size_t s = 0; s < n; ++s) p.heff[s] = v_add(p.heff[s], p.hs[s]);
    std::fill(p.hs.begin(), p.hs.end(), V3{{0.0, 0.0, 0.0}});
    cubic_add(p, m, p.hs.data());
    r->e_cub = energy_field_form(p, m, p.hs.data(), -0.5);
    for (size_t s = 0; s < n; ++s) p.heff[s] = v_add(p.heff[s], p.hs[s]);
    std::fill(p.hs.begin(), p.hs.end(), V3{{0.0, 0.0, 0.0}});
    dmi_add(p, m, p.hs.data());
    r->e_dmi = energy_field_form(p, m, p.hs.data(), -0.5);
    for (size_t s = 0; s < n; ++s) p.heff[s] = v_add(p.heff[s], p.hs[s]);
    mel_add(p, m, p.heff.data());
    thermal_add(p, p.heff.data());
    oe_add(p, p.heff.data());
    r->e_mel = 0.0;
    if (p.d.has_mel) {
        const double vol = p.d.dx * p.d.dy * p.d.dz;
        for (size_t s = 0; s < n; ++s)
            if (p.active(s)) r->e_mel += mel_density(p, m[s]) * vol * p.phi_at(s);
    }
    r->max_heff = max_norm(p.heff.data(), n);
    for (size_t s = 0; s < n; ++s) p.rhs[s] = llg(p, m[s], p.heff[s]);
    torque_add(p, m, p.rhs.data());
    r->max_rhs = max_norm(p.rhs.data(), n);
    r->e_total = r->e_ex + r->e_demag + r->e_ext + r->e_ani + r->e_cub + r->e_dmi + r->e_mel;
}

}  // namespace

extern "C" {

double orc_newell_f(double x, double y, double z) { return newell_f(x, y, z); }
double orc_newell_g(double x, double y, double z) { return newell_g(x, y, z); }

// out: 6 * (2nx*2ny*2nz) doubles, order xx,yy,zz,xy,xz,yz.
int orc_newell_kernels(uint32_t nx, uint32_t ny, uint32_t nz, double dx, double dy, double dz, double *out) {
    if (!nx || !ny || !nz || !out) return -1;
    newell_kernels((int)nx, (int)ny, (int)nz, dx, dy, dz, out);
    return 0;
}

// In-place complex FFT of `n` interleaved (re,im) doubles, unnormalised.
int orc_fft(double *data, int n, int inverse) {
    if (n < 1 || !data) return -1;
    FftPlan plan(n);
    std::vector<cplx> scr(2 * (size_t)n);
    plan.run(reinterpret_cast<cplx *>(data), 1, inverse != 0, scr.data());
    return 0;
}

// One step of the planner's xorshift64 (util.rs:128-130: shifts 13, 7, 17,
// left-right-left, Marsaglia's xor64).
static inline uint64_t xorshift64_next(uint64_t &st) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
}

// The first n states after `seed` (the values util.rs maps to components).
int orc_xorshift64_states(uint64_t seed, uint64_t n, uint64_t *out) {
    if (!out) return -1;
    uint64_t st = seed;
    for (uint64_t i = 0; i < n; ++i) out[i] = xorshift64_next(st);
    return 0;
}

// util.rs:133: component in [-1, 1] of a state (Rust `as f64` rounds to
// nearest, and u64::MAX as f64 is 2^64).
double orc_state_component(uint64_t st) { return ((double)st / (double)UINT64_MAX) * 2.0 - 1.0; }

// util.rs:118-147.  Seed 0 never leaves the rejection loop, so it is refused.
int orc_random_unit_vectors(uint64_t seed, uint64_t count, double *out) {
    if (seed == 0 || !out) return -1;
    uint64_t st = seed;
    for (uint64_t i = 0; i < count; ++i) {
        double c[3];
        for (;;) {
            for (int k = 0; k < 3; ++k) c[k] = orc_state_component(xorshift64_next(st));
            const double nrm = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            if (nrm > 1e-10) {
                out[3 * i + 0] = c[0] / nrm;
                out[3 * i + 1] = c[1] / nrm;
                out[3 * i + 2] = c[2] / nrm;
                break;
            }
        }
    }
    return 0;
}

// Box minus centred cylinder, geometry.rs:159-231 cell-centre rule.
uint64_t orc_hole_mask(uint32_t nx, uint32_t ny, uint32_t nz, double dx, double dy, double radius, uint8_t *mask) {
    const double cx = nx * dx * 0.5, cy = ny * dy * 0.5, r2 = radius * radius;
    uint64_t act = 0;
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t x = 0; x < nx; ++x) {
                const double px = (x + 0.5) * dx - cx, py = (y + 0.5) * dy - cy;
                const size_t i = x + (size_t)nx * (y + (size_t)ny * z);
                mask[i] = (px * px + py * py) <= r2 ? 0 : 1;
                act += mask[i];
            }
    return act;
}

void *orc_create(const orc_desc *d) {
    if (!d || !d->nx || !d->ny || !d->nz) return nullptr;
    Problem *p = new Problem();
    p->d = *d;
    p->n = (size_t)d->nx * d->ny * d->nz;
    p->has_mask = d->mask != nullptr;
    if (p->has_mask) p->mask.assign(d->mask, d->mask + p->n);
    p->d.mask = nullptr;
    if (d->ms_cell) p->ms_c.assign(d->ms_cell, d->ms_cell + p->n);
    if (d->ku1_cell) p->ku1_c.assign(d->ku1_cell, d->ku1_cell + p->n);
    if (d->ku2_cell) p->ku2_c.assign(d->ku2_cell, d->ku2_cell + p->n);
    if (d->axis_cell) p->axis_c.assign(d->axis_cell, d->axis_cell + 3 * p->n);
    if (d->wave_mask) p->wmask.assign(d->wave_mask, d->wave_mask + p->n);
    if (d->kc1_cell) p->kc1_c.assign(d->kc1_cell, d->kc1_cell + p->n);
    if (d->kc2_cell) p->kc2_c.assign(d->kc2_cell, d->kc2_cell + p->n);
    if (d->kc3_cell) p->kc3_c.assign(d->kc3_cell, d->kc3_cell + p->n);
    p->d.ms_cell = p->d.ku1_cell = p->d.ku2_cell = p->d.axis_cell = nullptr;
    p->d.kc1_cell = p->d.kc2_cell = p->d.kc3_cell = nullptr;
    p->d.wave_mask = nullptr;
    if (d->has_oe && d->oe_radius > 0.0) oe_profile(*p);
    if (d->bc_tier > 0 && d->bc_phi) {
        p->phi.assign(d->bc_phi, d->bc_phi + p->n);
        for (int q = 0; q < 6; ++q) {
            if (d->bc_face[0] && d->bc_face[q]) p->face[q].assign(d->bc_face[q], d->bc_face[q] + p->n);
            if (d->bc_delta[0] && d->bc_delta[q]) p->dlt[q].assign(d->bc_delta[q], d->bc_delta[q] + p->n);
        }
        if (d->dc_count && d->dc_stencil && d->dc_target && d->dc_source && d->dc_tensor) {
            const size_t ns = (size_t)d->dc_count * d->dc_stencil;
            p->dc_t.assign(d->dc_target, d->dc_target + d->dc_count);
            p->dc_s.assign(d->dc_source, d->dc_source + ns);
            p->dc_n.assign(d->dc_tensor, d->dc_tensor + 6 * ns);
        }
    }
    p->d.bc_phi = nullptr;
    for (int q = 0; q < 6; ++q) p->d.bc_face[q] = p->d.bc_delta[q] = nullptr;
    p->d.dc_target = p->d.dc_source = nullptr;
    p->d.dc_tensor = nullptr;
    const size_t n = p->n;
    p->m0.resize(n);
    p->ms.resize(n);
    p->heff.resize(n);
    p->hs.resize(n);
    p->rhs.resize(n);
    for (auto &k : p->k) k.resize(n);
    if (d->layer_z) p->lz.assign(d->layer_z, d->layer_z + d->nz);
    p->d.layer_z = nullptr;
    if (d->region && d->exchange_lut) {
        p->reg.assign(d->region, d->region + n);
        for (uint32_t r : p->reg)
            if (r >= 256) {
                delete p;
                return nullptr;
            }
        p->lut.assign(d->exchange_lut, d->exchange_lut + 256 * 256);
    }
    p->d.region = nullptr;
    p->d.exchange_lut = nullptr;
    if (d->demag && p->layered()) {
        p->px = 2 * d->nx;
        p->py = 2 * d->ny;
        p->pz = 2;
        p->P = (size_t)p->px * p->py * p->pz;
        p->fx = FftPlan(p->px);
        p->fy = FftPlan(p->py);
        p->fz = FftPlan(p->pz);
        layer_kernels(*p);
    } else if (d->demag) {
        p->px = p->per(0) ? d->nx : 2 * d->nx;
        p->py = p->per(1) ? d->ny : 2 * d->ny;
        p->pz = p->per(2) ? d->nz : 2 * d->nz;
        p->P = (size_t)p->px * p->py * p->pz;
        p->fx = FftPlan(p->px);
        p->fy = FftPlan(p->py);
        p->fz = FftPlan(p->pz);
        std::vector<double> nk(6 * p->P);
        if (d->periodic) periodic_kernels(*p, nk.data());
        else newell_kernels(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, nk.data());
        std::vector<cplx> *dst[6] = {&p->kxx, &p->kyy, &p->kzz, &p->kxy, &p->kxz, &p->kyz};
        for (int c = 0; c < 6; ++c) {
            dst[c]->resize(p->P);
            for (size_t i = 0; i < p->P; ++i) (*dst[c])[i] = cplx(nk[c * p->P + i], 0.0);
            fft3(dst[c]->data(), p->px, p->py, p->pz, p->fx, p->fy, p->fz, false);
        }
        for (auto *b : {&p->bmx, &p->bmy, &p->bmz, &p->bhx, &p->bhy, &p->bhz}) b->assign(p->P, cplx(0.0));
    }
    return p;
}

void orc_destroy(void *h) { delete static_cast<Problem *>(h); }

// Kernel spectra as used by the reference GPU ABI (interleaved re/im over the
// full (2nx,2ny,2nz) grid), component c in xx,yy,zz,xy,xz,yz order.
int orc_kernel_spectrum(void *h, int c, double *out) {
    Problem *p = static_cast<Problem *>(h);
    if (!p || !p->d.demag || c < 0 || c > 5) return -1;
    const std::vector<cplx> *src[6] = {&p->kxx, &p->kyy, &p->kzz, &p->kxy, &p->kxz, &p->kyz};
    std::memcpy(out, src[c]->data(), sizeof(cplx) * p->P);
    return 0;
}

// Individual field terms at magnetisation m (AoS 3n).  Any output may be NULL.
int orc_fields(void *h, const double *m, double *h_ex, double *h_d, double *h_ext, double *h_ani, double *h_eff) {
    Problem *p = static_cast<Problem *>(h);
    const V3 *mm = reinterpret_cast<const V3 *>(m);
    std::vector<V3> t(p->n);
    auto one = [&](double *out, int which) {
        if (!out) return;
        std::fill(t.begin(), t.end(), V3{{0.0, 0.0, 0.0}});
        if (which == 0 && p->d.exchange) exchange_add(*p, mm, t.data());
        if (which == 1 && p->d.demag) demag_add(*p, mm, t.data());
        if (which == 2) ext_add(*p, t.data());
        if (which == 3) uni_add(*p, mm, t.data());
        std::memcpy(out, t.data(), sizeof(V3) * p->n);
    };
    one(h_ex, 0);
    one(h_d, 1);
    one(h_ext, 2);
    one(h_ani, 3);
    if (h_eff) {
        effective_field(*p, mm, p->heff.data());
        std::memcpy(h_eff, p->heff.data(), sizeof(V3) * p->n);
    }
    return 0;
}

// Brute-force real-space demag: H_i = -sum_j N(r_i - r_j) . Ms m_j, using the
// same Newell tensor.  O(n^2); for pinning the FFT path on small grids.
int orc_demag_direct(void *h, const double *m, double *h_d) {
    Problem *p = static_cast<Problem *>(h);
    const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz;
    const int px = 2 * nx, py = 2 * ny, pz = 2 * nz;
    const size_t P = (size_t)px * py * pz;
    std::vector<double> nk(6 * P);
    newell_kernels(nx, ny, nz, p->d.dx, p->d.dy, p->d.dz, nk.data());
    const double *N = nk.data();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)p->n; ++i) {
        double hx = 0, hy = 0, hz = 0;
        if (p->active(i)) {
            const int xi = i % nx, yi = (i / nx) % ny, zi = i / (nx * ny);
            for (size_t j = 0; j < p->n; ++j) {
                if (!p->active(j)) continue;
                const int xj = j % nx, yj = (j / nx) % ny, zj = j / (nx * ny);
                const int ax = (xi - xj + px) % px, ay = (yi - yj + py) % py, az = (zi - zj + pz) % pz;
                const size_t q = ((size_t)az * py + ay) * px + ax;
                const double Mx = m[3 * j] * p->d.ms, My = m[3 * j + 1] * p->d.ms, Mz = m[3 * j + 2] * p->d.ms;
                hx -= N[q] * Mx + N[3 * P + q] * My + N[4 * P + q] * Mz;
                hy -= N[3 * P + q] * Mx + N[P + q] * My + N[5 * P + q] * Mz;
                hz -= N[4 * P + q] * Mx + N[5 * P + q] * My + N[2 * P + q] * Mz;
            }
        }
        h_d[3 * i] = hx;
        h_d[3 * i + 1] = hy;
        h_d[3 * i + 2] = hz;
    }
    return 0;
}

int orc_rhs(void *h, const double *m, double *k) {
    Problem *p = static_cast<Problem *>(h);
    rhs_eval(*p, reinterpret_cast<const V3 *>(m), reinterpret_cast<V3 *>(k));
    return 0;
}

int orc_observables(void *h, const double *m, orc_report *r) {
    Problem *p = static_cast<Problem *>(h);
    observables(*p, reinterpret_cast<const V3 *>(m), r);
    return 0;
}

// Normalise (ExchangeLlgState::new) and zero inactive cells (new_state).
int orc_prepare_state(void *h, double *m) {
    Problem *p = static_cast<Problem *>(h);
    V3 *mm = reinterpret_cast<V3 *>(m);
    for (size_t s = 0; s < p->n; ++s) mm[s] = p->active(s) ? v_normalized(mm[s]) : V3{{0.0, 0.0, 0.0}};
    p->fsal_valid = false;
    p->abm_count = 0;
    p->abm_dt = 0.0;
    return 0;
}

void orc_invalidate_fsal(void *h) { static_cast<Problem *>(h)->fsal_valid = false; }

// ExchangeLlgState::reset_abm_history (fdm_state.rs:172-174).
void orc_reset_abm(void *h) {
    Problem *p = static_cast<Problem *>(h);
    p->abm_count = 0;
    p->abm_dt = 0.0;
}

int orc_abm_count(void *h) { return static_cast<Problem *>(h)->abm_count; }

// Time at which orc_fields / orc_rhs / orc_observables evaluate the waveform.
void orc_set_time(void *h, double t) { static_cast<Problem *>(h)->t_eval = t; }

// fdm_integrators.rs:84-150
static int heun_impl(void *h, double *m, double *t, double dt, orc_report *rep) {
    Problem *p = static_cast<Problem *>(h);
    p->th_dt_cur = dt;
    const size_t n = p->n;
    V3 *mm = reinterpret_cast<V3 *>(m);
    std::copy(mm, mm + n, p->m0.begin());
    const double t0 = *t;
    p->t_eval = t0;
    rhs_eval(*p, p->m0.data(), p->k[0].data());
    for (size_t s = 0; s < n; ++s) p->ms[s] = v_normalized(v_add(p->m0[s], v_scale(p->k[0][s], dt)));
    p->t_eval = t0 + dt;
    rhs_eval(*p, p->ms.data(), p->k[1].data());
    for (size_t s = 0; s < n; ++s) mm[s] = v_normalized(v_add(p->m0[s], v_scale(v_add(p->k[0][s], p->k[1][s]), 0.5 * dt)));
    *t += dt;
    p->t_eval = *t;
    observables(*p, mm, rep);
    rep->time = *t;
    rep->dt_used = dt;
    rep->dt_next = 0.0;
    rep->rejected = 0;
    p->fsal_valid = false;
    return 0;
}

// fdm_integrators.rs:155-264
static int rk4_impl(void *h, double *m, double *t, double dt, orc_report *rep) {
    Problem *p = static_cast<Problem *>(h);
    p->th_dt_cur = dt;
    const size_t n = p->n;
    V3 *mm = reinterpret_cast<V3 *>(m);
    std::copy(mm, mm + n, p->m0.begin());
    auto &K = p->k;
    const double t0 = *t;
    p->t_eval = t0;
    rhs_eval(*p, p->m0.data(), K[0].data());
    for (size_t s = 0; s < n; ++s) p->ms[s] = v_normalized(v_add(p->m0[s], v_scale(K[0][s], 0.5 * dt)));
    p->t_eval = t0 + 0.5 * dt;
    rhs_eval(*p, p->ms.data(), K[1].data());
    for (size_t s = 0; s < n; ++s) p->ms[s] = v_normalized(v_add(p->m0[s], v_scale(K[1][s], 0.5 * dt)));
    rhs_eval(*p, p->ms.data(), K[2].data());
    for (size_t s = 0; s < n; ++s) p->ms[s] = v_normalized(v_add(p->m0[s], v_scale(K[2][s], dt)));
    p->t_eval = t0 + dt;
    rhs_eval(*p, p->ms.data(), K[3].data());
    const double dt6 = dt / 6.0;
    for (size_t s = 0; s < n; ++s)
        mm[s] = v_normalized(v_add(
            p->m0[s], v_scale(v_add(v_add(K[0][s], v_scale(K[1][s], 2.0)), v_add(v_scale(K[2][s], 2.0), K[3][s])), dt6)));
    *t += dt;
    observables(*p, mm, rep);
    rep->time = *t;
    rep->dt_used = dt;
    rep->dt_next = 0.0;
    rep->rejected = 0;
    p->fsal_valid = false;
    return 0;
}

// Dormand-Prince 5(4) with FSAL, fdm_integrators.rs:429-680,828-870.
static int rk45_impl(void *h, double *m, double *t, double dt_in, orc_report *rep) {
    Problem *p = static_cast<Problem *>(h);
    const orc_desc &c = p->d;
    const size_t n = p->n;
    V3 *mm = reinterpret_cast<V3 *>(m);
    double dt = std::max(std::min(dt_in, c.dt_max), c.dt_min);
    std::copy(mm, mm + n, p->m0.begin());
    const double A21 = 1.0 / 5.0;
    const double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
    const double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
    const double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0, A54 = -212.0 / 729.0;
    const double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0, A64 = 49.0 / 176.0,
                 A65 = -5103.0 / 18656.0;
    const double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0, B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
    const double E1 = 71.0 / 57600.0, E3 = -71.0 / 16695.0, E4 = 71.0 / 1920.0, E5 = -17253.0 / 339200.0,
                 E6 = 22.0 / 525.0, E7 = -1.0 / 40.0;
    auto &K = p->k;
    const V3 *m0 = p->m0.data();
    V3 *st = p->ms.data();
    uint64_t rejected = 0;
    const double t0 = *t;
    for (;;) {
        p->th_dt_cur = dt;
        p->t_eval = t0;
        if (p->fsal_valid) {
            std::copy(p->kfsal.begin(), p->kfsal.end(), K[0].begin());
            p->fsal_valid = false;
        } else {
            rhs_eval(*p, m0, K[0].data());
        }
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s) st[s] = v_normalized(v_add(m0[s], v_scale(K[0][s], A21 * dt)));
        p->t_eval = t0 + (1.0 / 5.0) * dt;
        rhs_eval(*p, st, K[1].data());
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s)
            st[s] = v_normalized(v_add(m0[s], v_scale(v_add(v_scale(K[0][s], A31), v_scale(K[1][s], A32)), dt)));
        p->t_eval = t0 + (3.0 / 10.0) * dt;
        rhs_eval(*p, st, K[2].data());
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s)
            st[s] = v_normalized(v_add(
                m0[s], v_scale(v_add(v_add(v_scale(K[0][s], A41), v_scale(K[1][s], A42)), v_scale(K[2][s], A43)), dt)));
        p->t_eval = t0 + (4.0 / 5.0) * dt;
        rhs_eval(*p, st, K[3].data());
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s)
            st[s] = v_normalized(v_add(m0[s], v_scale(v_add(v_add(v_scale(K[0][s], A51), v_scale(K[1][s], A52)),
                                                            v_add(v_scale(K[2][s], A53), v_scale(K[3][s], A54))),
                                                      dt)));
        p->t_eval = t0 + (8.0 / 9.0) * dt;
        rhs_eval(*p, st, K[4].data());
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s)
            st[s] = v_normalized(v_add(
                m0[s],
                v_scale(v_add(v_add(v_add(v_scale(K[0][s], A61), v_scale(K[1][s], A62)), v_scale(K[2][s], A63)),
                              v_add(v_scale(K[3][s], A64), v_scale(K[4][s], A65))),
                        dt)));
        p->t_eval = t0 + 1.0 * dt;
        rhs_eval(*p, st, K[5].data());
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s)
            st[s] = v_normalized(v_add(
                m0[s], v_scale(v_add(v_add(v_add(v_scale(K[0][s], B1), v_scale(K[2][s], B3)), v_scale(K[3][s], B4)),
                                     v_add(v_scale(K[4][s], B5), v_scale(K[5][s], B6))),
                               dt)));
        p->t_eval = t0 + 1.0 * dt;
        rhs_eval(*p, st, K[6].data());
        // error norm
        const bool use_rtol = c.rtol > 0.0;
        double err = 0.0;
        const int idx[6] = {0, 2, 3, 4, 5, 6};
        const double w[6] = {E1, E3, E4, E5, E6, E7};
#pragma omp parallel for reduction(max : err) schedule(static)
        for (long s = 0; s < (long)n; ++s) {
            double e[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < 6; ++q)
                for (int a = 0; a < 3; ++a) e[a] += w[q] * K[idx[q]][s].c[a];
            for (int a = 0; a < 3; ++a) e[a] *= dt;
            double v = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            if (use_rtol) v /= (c.atol + c.rtol * std::max(v_norm(m0[s]), 1e-30));
            err = std::max(err, v);
        }
        const double thr = use_rtol ? 1.0 : c.atol;
        if (err <= thr || dt <= c.dt_min) {
            std::copy(st, st + n, mm);
            *t += dt;
            p->kfsal = K[6];
            p->fsal_valid = true;
            const double ratio = std::max(std::min(c.headroom * std::pow(thr / std::max(err, 1e-30), 0.2), c.growth), c.shrink);
            const double dt_next = std::min(std::max(dt * ratio, c.dt_min), c.dt_max);
            p->t_eval = *t;
            observables(*p, mm, rep);
            rep->time = *t;
            rep->dt_used = dt;
            rep->dt_next = dt_next;
            rep->rejected = rejected;
            return 0;
        }
        ++rejected;
        const double ratio = std::max(std::min(c.headroom * std::pow(thr / err, 0.2), c.growth), c.shrink);
        dt = std::min(std::max(dt * ratio, c.dt_min), c.dt_max);
    }
}

// Bogacki-Shampine 3(2), fdm_integrators.rs:290-410: k1 is evaluated on
// every attempt (no FSAL), stages normalised, 3rd-order solution, error
// dt*(-5/72 k1 + 1/12 k2 + 1/9 k3 - 1/8 k4) with k4 at the new state, and
// the controller exponent 1/3.  The waveform is evaluated at the stage
// times t0, t0 + dt/2, t0 + 3dt/4, t0 + dt.
static int rk23_impl(void *h, double *m, double *t, double dt_in, orc_report *rep) {
    Problem *p = static_cast<Problem *>(h);
    const orc_desc &c = p->d;
    const size_t n = p->n;
    V3 *mm = reinterpret_cast<V3 *>(m);
    double dt = std::max(std::min(dt_in, c.dt_max), c.dt_min);
    std::copy(mm, mm + n, p->m0.begin());
    auto &K = p->k;
    const V3 *m0 = p->m0.data();
    V3 *st = p->ms.data();
    uint64_t rejected = 0;
    const double t0 = *t;
    for (;;) {
        p->th_dt_cur = dt;
        p->t_eval = t0;
        rhs_eval(*p, m0, K[0].data());
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s) st[s] = v_normalized(v_add(m0[s], v_scale(K[0][s], 0.5 * dt)));
        p->t_eval = t0 + 0.5 * dt;
        rhs_eval(*p, st, K[1].data());
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s) st[s] = v_normalized(v_add(m0[s], v_scale(K[1][s], 0.75 * dt)));
        p->t_eval = t0 + 0.75 * dt;
        rhs_eval(*p, st, K[2].data());
#pragma omp parallel for schedule(static)
        for (long s = 0; s < (long)n; ++s)
            st[s] = v_normalized(v_add(
                m0[s], v_scale(v_add(v_add(v_scale(K[0][s], 2.0 / 9.0), v_scale(K[1][s], 1.0 / 3.0)),
                                     v_scale(K[2][s], 4.0 / 9.0)),
                               dt)));
        p->t_eval = t0 + dt;
        rhs_eval(*p, st, K[3].data());
        const bool use_rtol = c.rtol > 0.0;
        double err = 0.0;
        const double w[4] = {-5.0 / 72.0, 1.0 / 12.0, 1.0 / 9.0, -1.0 / 8.0};
#pragma omp parallel for reduction(max : err) schedule(static)
        for (long s = 0; s < (long)n; ++s) {
            double e[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < 4; ++q)
                for (int a = 0; a < 3; ++a) e[a] += w[q] * K[q][s].c[a];
            for (int a = 0; a < 3; ++a) e[a] *= dt;
            double v = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            if (use_rtol) v /= (c.atol + c.rtol * std::max(v_norm(m0[s]), 1e-30));
            err = std::max(err, v);
        }
        const double thr = use_rtol ? 1.0 : c.atol;
        if (err <= thr || dt <= c.dt_min) {
            std::copy(st, st + n, mm);
            *t += dt;
            const double ratio =
                std::max(std::min(c.headroom * std::pow(thr / std::max(err, 1e-30), 1.0 / 3.0), c.growth), c.shrink);
            const double dt_next = std::min(std::max(dt * ratio, c.dt_min), c.dt_max);
            p->t_eval = *t;
            observables(*p, mm, rep);
            rep->time = *t;
            rep->dt_used = dt;
            rep->dt_next = dt_next;
            rep->rejected = rejected;
            p->fsal_valid = false;
            return 0;
        }
        ++rejected;
        const double ratio = std::max(std::min(c.headroom * std::pow(thr / err, 1.0 / 3.0), c.growth), c.shrink);
        dt = std::min(std::max(dt * ratio, c.dt_min), c.dt_max);
    }
}

// AbmHistory::push (fdm_state.rs:247-257): a dt that differs from the last
// push by more than 10% restarts the history first; then f becomes f_n.
static void abm_push(Problem &p, std::vector<V3> &f, double dt) {
    if (p.abm_dt > 0.0 && std::fabs(dt - p.abm_dt) / p.abm_dt > 0.1) {
        p.abm_count = 0;
        p.abm_dt = 0.0;
    }
    std::swap(p.abm_f[2], p.abm_f[1]);
    std::swap(p.abm_f[1], p.abm_f[0]);
    std::swap(p.abm_f[0], f);
    p.abm_count = std::min(p.abm_count + 1, 3);
    p.abm_dt = dt;
}

// fdm_integrators.rs:683-820.  Until three RHS values are in the history
// the step is a Heun step and pushes f(m_new); afterwards the
// Adams-Bashforth predictor, f* = f(pred) and the Adams-Moulton corrector,
// pushing f*.
static int abm3_impl(void *h, double *m, double *t, double dt, orc_report *rep) {
    Problem *p = static_cast<Problem *>(h);
    p->th_dt_cur = dt;
    const size_t n = p->n;
    V3 *mm = reinterpret_cast<V3 *>(m);
    std::vector<V3> f(n);
    if (p->abm_count < 3) {
        heun_impl(p, m, t, dt, rep);
        p->t_eval = *t;
        rhs_eval(*p, mm, f.data());
        abm_push(*p, f, dt);
        return 0;
    }
    std::copy(mm, mm + n, p->m0.begin());
    const V3 *fn = p->abm_f[0].data(), *fn1 = p->abm_f[1].data(), *fn2 = p->abm_f[2].data();
    const double t0 = *t;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)n; ++s) {
        const V3 pr = v_add(v_add(v_scale(fn[s], 23.0 / 12.0), v_scale(fn1[s], -16.0 / 12.0)), v_scale(fn2[s], 5.0 / 12.0));
        p->ms[s] = v_normalized(v_add(p->m0[s], v_scale(pr, dt)));
    }
    p->t_eval = t0 + dt;
    rhs_eval(*p, p->ms.data(), f.data());
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)n; ++s) {
        const V3 co = v_add(v_add(v_scale(f[s], 5.0 / 12.0), v_scale(fn[s], 8.0 / 12.0)), v_scale(fn1[s], -1.0 / 12.0));
        mm[s] = v_normalized(v_add(p->m0[s], v_scale(co, dt)));
    }
    *t += dt;
    abm_push(*p, f, dt);
    p->t_eval = *t;
    observables(*p, mm, rep);
    rep->time = *t;
    rep->dt_used = dt;
    rep->dt_next = 0.0;
    rep->rejected = 0;
    p->fsal_valid = false;
    return 0;
}

// One step call each; the thermal counter advances after every call
// (FdmProblem::step_with_buffers, fdm_problem.rs:225-247).
int orc_heun_step(void *h, double *m, double *t, double dt, orc_report *rep) {
    const int r = heun_impl(h, m, t, dt, rep);
    ++static_cast<Problem *>(h)->th_step;
    return r;
}
int orc_rk4_step(void *h, double *m, double *t, double dt, orc_report *rep) {
    const int r = rk4_impl(h, m, t, dt, rep);
    ++static_cast<Problem *>(h)->th_step;
    return r;
}
int orc_rk45_step(void *h, double *m, double *t, double dt, orc_report *rep) {
    const int r = rk45_impl(h, m, t, dt, rep);
    ++static_cast<Problem *>(h)->th_step;
    return r;
}
int orc_rk23_step(void *h, double *m, double *t, double dt, orc_report *rep) {
    const int r = rk23_impl(h, m, t, dt, rep);
    ++static_cast<Problem *>(h)->th_step;
    return r;
}
int orc_abm3_step(void *h, double *m, double *t, double dt, orc_report *rep) {
    const int r = abm3_impl(h, m, t, dt, rep);
    ++static_cast<Problem *>(h)->th_step;
    return r;
}

uint64_t orc_thermal_step(void *h) { return static_cast<Problem *>(h)->th_step; }

// SplitMix64 finaliser (for pinning against its published outputs).
uint64_t orc_splitmix64(uint64_t z) { return splitmix64(z); }

// Runner loop, cpu_reference.rs:622-639: dt_step = min(dt, until - t); the
// next dt is the integrator's suggestion.  integrator: 1 Heun, 2 DP45, 3 ABM3,
// 4 RK4, 5 RK23.
// Records each accepted step's (time, dt) in `trace` (2 doubles per step) if
// non-NULL.  Returns the number of steps taken.
uint64_t orc_run(void *h, int integrator, double *m, double *t, double until, double dt0, uint64_t max_steps,
                 orc_report *last, double *trace) {
    Problem *p = static_cast<Problem *>(h);
    double dt = dt0;
    uint64_t steps = 0;
    orc_report rep{};
    while (*t < until && steps < max_steps) {
        const double dt_step = std::min(dt, until - *t);
        if (integrator == 2) {
            orc_rk45_step(p, m, t, dt_step, &rep);
            dt = rep.dt_next;
        } else if (integrator == 5) {
            orc_rk23_step(p, m, t, dt_step, &rep);
            dt = rep.dt_next;
        } else if (integrator == 4) {
            orc_rk4_step(p, m, t, dt_step, &rep);
        } else if (integrator == 3) {
            orc_abm3_step(p, m, t, dt_step, &rep);
        } else {
            orc_heun_step(p, m, t, dt_step, &rep);
        }
        if (trace) {
            trace[2 * steps] = rep.time;
            trace[2 * steps + 1] = rep.dt_used;
        }
        ++steps;
    }
    if (last) *last = rep;
    return steps;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
