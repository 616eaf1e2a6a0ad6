// newell_host.cpp -- FP64 Newell demagnetisation tensor and its
// symmetry-reduced real spectrum, computed once at create().
//
// Tensor: Newell, Williams & Dunlop (1993) f/g potentials, 27-point second
// difference with Neumaier-compensated summation, point-dipole far field
// beyond 40 cells, octant parity (xx,yy,zz even; xy odd in x,y; xz odd in
// x,z; yz odd in y,z).  This is the algorithm of the reference's
// crates/fullmag-fdm-demag/src/newell.rs:21-517; the formulas, the 40-cell
// threshold and the stencil weights must be identical for the tensor to be
// the same operator.
//
// Spectrum: the reference transforms the full zero-padded real tensor with
// complex FFTs (crates/fullmag-engine/src/fdm_fft.rs:334-432).  Because each
// component has definite parity along every axis its spectrum is real and
// has the same parity in k, so only kx in [0,nx], ky in [0,ny], kz in [0,nz]
// is kept: thin film 4 x (nx+1)(ny+1) values, 3-D 6 x (nx+1)(ny+1)(nz+1).
// Transform: per axis, a complex FFT of each (parity-extended) line of the
// first-octant data, keeping the lower half of the output.
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>
#include <functional>

namespace fm {
namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;
constexpr int kFar = 40;

double pot_f(double x, double y, double z) {
    const double x2 = x * x, y2 = y * y, z2 = z * z, r2 = x2 + y2 + z2;
    if (r2 < 1e-300) return 0.0;
    const double r = std::sqrt(r2);
    double v = (2.0 * x2 - y2 - z2) * r / 6.0;
    if (x2 + z2 > 1e-300) {
        const double u = 2.0 * y * (y + r) / (x2 + z2);
        if (u > -1.0) v += y * (z2 - x2) / 4.0 * std::log1p(u);
    }
    if (x2 + y2 > 1e-300) {
        const double u = 2.0 * z * (z + r) / (x2 + y2);
        if (u > -1.0) v += z * (y2 - x2) / 4.0 * std::log1p(u);
    }
    if (std::fabs(x) > 1e-300) v -= x * y * z * std::atan(y * z / (x * r));
    return v;
}

double pot_g(double x, double y, double z) {
    const double x2 = x * x, y2 = y * y, z2 = z * z, r2 = x2 + y2 + z2;
    if (r2 < 1e-300) return 0.0;
    const double r = std::sqrt(r2);
    double v = -x * y * r / 3.0;
    if (x2 + y2 > 1e-300) {
        const double u = 2.0 * z * (z + r) / (x2 + y2);
        if (u > -1.0) v += x * y * z * std::log1p(u) / 2.0;
    }
    if (y2 + z2 > 1e-300) {
        const double u = 2.0 * x * (x + r) / (y2 + z2);
        if (u > -1.0) v += y * (3.0 * z2 - y2) * std::log1p(u) / 12.0;
    }
    if (x2 + z2 > 1e-300) {
        const double u = 2.0 * y * (y + r) / (x2 + z2);
        if (u > -1.0) v += x * (3.0 * z2 - x2) * std::log1p(u) / 12.0;
    }
    if (std::fabs(z) > 1e-300) v -= z2 * z / 6.0 * std::atan(x * y / (z * r));
    if (std::fabs(y) > 1e-300) v -= y2 * z / 2.0 * std::atan(x * z / (y * r));
    if (std::fabs(x) > 1e-300) v -= x2 * z / 2.0 * std::atan(y * z / (x * r));
    return v;
}

void parallel_for(long n, const std::function<void(long, long)> &body) {
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 64 || nt == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    const long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const long a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back(body, a, b);
    }
    for (auto &t : th) t.join();
}

// Compensated 27-point sum: weights 8 (centre), -4 (faces), 2 (edges),
// -1 (corners) of the potential sampled at (i+a, j+b, k+c), a,b,c in {-1,0,1}.
double second_difference(const std::vector<double> &pv, int sx, int sy, int i, int j, int k) {
    double s = 0.0, comp = 0.0;
    for (int c = -1; c <= 1; ++c)
        for (int b = -1; b <= 1; ++b)
            for (int a = -1; a <= 1; ++a) {
                const int nzc = (a != 0) + (b != 0) + (c != 0);
                const double w = nzc == 0 ? 8.0 : nzc == 1 ? -4.0 : nzc == 2 ? 2.0 : -1.0;
                const double t = w * pv[(size_t)(k + c + 1) * sy * sx + (size_t)(j + b + 1) * sx + (i + a + 1)];
                const double u = s + t;
                comp += std::fabs(s) >= std::fabs(t) ? (s - u) + t : (t - u) + s;
                s = u;
            }
    return s + comp;
}

// Mixed-radix Stockham FFT (forward, unnormalised) on one complex line.
struct LineFft {
    int n;
    std::vector<int> rad;
    std::vector<cd> w;
    explicit LineFft(int n_) : n(n_) {
        int m = n;
        while (m % 4 == 0) { rad.push_back(4); m /= 4; }
        while (m % 2 == 0) { rad.push_back(2); m /= 2; }
        for (int p = 3; m > 1; p += 2)
            while (m % p == 0) { rad.push_back(p); m /= p; }
        w.resize(n);
        for (int k = 0; k < n; ++k) w[k] = std::polar(1.0, -2.0 * kPi * (double)k / (double)n);
    }
    void run(cd *a, cd *b) const {  // result in a; b is scratch of n
        int ns = 1;
        std::vector<cd> v, o;
        for (int R : rad) {
            v.assign(R, 0.0);
            o.assign(R, 0.0);
            const int nr = n / R, ts = n / (ns * R);
            for (int j = 0; j < nr; ++j) {
                const int k = j % ns;
                for (int r = 0; r < R; ++r) v[r] = a[j + r * nr] * (r && ns > 1 ? w[(size_t)k * r * ts] : cd(1.0));
                for (int q = 0; q < R; ++q) {
                    cd acc = 0.0;
                    for (int r = 0; r < R; ++r) acc += v[r] * w[(size_t)((r * q) % R) * (n / R)];
                    o[q] = acc;
                }
                const int base = (j / ns) * ns * R + k;
                for (int r = 0; r < R; ++r) b[base + r * ns] = o[r];
            }
            std::copy(b, b + n, a);
            ns *= R;
        }
    }
};

// Transform along one axis: `in` has extents (e0 = the axis, e1, e2) with
// the axis fastest when stride==1, or generic strides; keeps outputs 0..n.
// data layout: index(a, b, c) = a*sa + b*sb + c*sc over extents (n, nb, nc).
void axis_transform(const std::vector<cd> &in, std::vector<cd> &out, int n, int nb, int nc, size_t sa, size_t sb,
                    size_t sc, size_t oa, size_t ob, size_t oc, int sign) {
    const int P = 2 * n;
    LineFft f(P);
    parallel_for((long)nb * nc, [&](long lo, long hi) {
        std::vector<cd> a(P), b(P);
        for (long l = lo; l < hi; ++l) {
            const long bi = l % nb, ci = l / nb;
            std::fill(a.begin(), a.end(), cd(0.0));
            for (int i = 0; i < n; ++i) {
                const cd v = in[(size_t)i * sa + (size_t)bi * sb + (size_t)ci * sc];
                a[i] = v;
                if (i > 0) a[P - i] = (double)sign * v;
            }
            f.run(a.data(), b.data());
            for (int k = 0; k <= n; ++k) out[(size_t)k * oa + (size_t)bi * ob + (size_t)ci * oc] = a[k];
        }
    });
}

}  // namespace

// First-octant real-space tensor, 6 arrays (xx,yy,zz,xy,xz,yz) of nx*ny*nz,
// index (k*ny + j)*nx + i = displacement (i dx, j dy, k dz).
std::vector<double> newell_octant(int nx, int ny, int nz, double dx, double dy, double dz) {
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> out(6 * n);
    const int fx = std::min(nx, kFar) + 2, fy = std::min(ny, kFar) + 2, fz = std::min(nz, kFar) + 2;
    const size_t fl = (size_t)fx * fy * fz;
    std::vector<double> pv[6];
    for (auto &v : pv) v.resize(fl);
    parallel_for(fz, [&](long lo, long hi) {
        for (long k = lo; k < hi; ++k)
            for (int j = 0; j < fy; ++j)
                for (int i = 0; i < fx; ++i) {
                    const double x = (i - 1) * dx, y = (j - 1) * dy, z = (k - 1) * dz;
                    const size_t id = (size_t)k * fy * fx + (size_t)j * fx + i;
                    pv[0][id] = pot_f(x, y, z);
                    pv[1][id] = pot_f(y, x, z);
                    pv[2][id] = pot_f(z, y, x);
                    pv[3][id] = pot_g(x, y, z);
                    pv[4][id] = pot_g(x, z, y);
                    pv[5][id] = pot_g(y, z, x);
                }
    });
    const double vol = dx * dy * dz;
    const double den = 4.0 * kPi * vol;
    parallel_for(nz, [&](long lo, long hi) {
        for (long k = lo; k < hi; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t o = ((size_t)k * ny + j) * nx + i;
                    const long d2 = (long)i * i + (long)j * j + k * k;
                    if (i >= kFar || j >= kFar || k >= kFar || d2 >= (long)kFar * kFar) {
                        const double x = i * dx, y = j * dy, z = k * dz;
                        const double r2 = x * x + y * y + z * z, r = std::sqrt(r2);
                        const double r3 = r2 * r, r5 = r2 * r2 * r;
                        auto dia = [&](double a) { return (1.0 / r3 - 3.0 * a * a / (r3 * r2)) / (4.0 * kPi) * vol; };
                        auto off = [&](double a, double b) { return -3.0 * a * b / (4.0 * kPi * r5) * vol; };
                        out[o] = dia(x);
                        out[n + o] = dia(y);
                        out[2 * n + o] = dia(z);
                        out[3 * n + o] = off(x, y);
                        out[4 * n + o] = off(x, z);
                        out[5 * n + o] = off(y, z);
                    } else {
                        for (int c = 0; c < 6; ++c) out[c * n + o] = second_difference(pv[c], fx, fy, i, j, (int)k) / den;
                    }
                }
    });
    return out;
}

// Symmetry-reduced real spectrum in the engine's layout:
//   thin film (nz == 1): [c][kx][ky], c in xx,yy,zz,xy
//   3-D:                [c][kx][ky][kz], c in xx,yy,zz,xy,xz,yz
std::vector<double> newell_reduced_spectrum(int nx, int ny, int nz, double dx, double dy, double dz) {
    const std::vector<double> oct = newell_octant(nx, ny, nz, dx, dy, dz);
    const size_t n = (size_t)nx * ny * nz;
    const bool film = nz == 1;
    const int ncomp = film ? 4 : 6;
    const int nkx = nx + 1, nky = ny + 1, nkz = film ? 1 : nz + 1;
    const size_t red = (size_t)nkx * nky * nkz;
    std::vector<double> out((size_t)ncomp * red);
    // parity (x,y,z) sign of the odd axes for each component
    const int par[6][3] = {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1}};
    for (int c = 0; c < ncomp; ++c) {
        // stage 0: octant (z, y, x) complex
        std::vector<cd> a(n);
        for (size_t i = 0; i < n; ++i) a[i] = oct[c * n + i];
        // x: (x, y, z) with strides (1, nx, nx*ny) -> (kx, y, z) strides (1, nkx, nkx*ny)
        std::vector<cd> b((size_t)nkx * ny * nz);
        axis_transform(a, b, nx, ny, nz, 1, nx, (size_t)nx * ny, 1, nkx, (size_t)nkx * ny, par[c][0]);
        // y: (y, kx, z) in b strides (nkx, 1, nkx*ny) -> (ky, kx, z) strides (nkx, 1, nkx*nky)
        std::vector<cd> d((size_t)nkx * nky * nz);
        axis_transform(b, d, ny, nkx, nz, nkx, 1, (size_t)nkx * ny, nkx, 1, (size_t)nkx * nky, par[c][1]);
        if (film) {
            for (int kx = 0; kx < nkx; ++kx)
                for (int ky = 0; ky < nky; ++ky) out[c * red + (size_t)kx * nky + ky] = d[(size_t)ky * nkx + kx].real();
        } else {
            // z: (z, kx, ky) in d strides (nkx*nky, 1, nkx) -> (kz, kx, ky)
            std::vector<cd> e((size_t)nkx * nky * nkz);
            axis_transform(d, e, nz, nkx, nky, (size_t)nkx * nky, 1, nkx, (size_t)nkx * nky, 1, nkx, par[c][2]);
            for (int kx = 0; kx < nkx; ++kx)
                for (int ky = 0; ky < nky; ++ky)
                    for (int kz = 0; kz < nkz; ++kz)
                        out[c * red + ((size_t)kx * nky + ky) * nkz + kz] =
                            e[((size_t)kz * nky + ky) * nkx + kx].real();
        }
    }
    return out;
}

// Thin-layer pair kernel of the multilayer runtime
// (crates/fullmag-fdm-demag/src/newell.rs:528-660, shifted_kernel.rs:24-46):
// the tensor between two single-cell-thick layers whose centres are `s`
// apart along z (s = z_destination - z_source) is N(i dx, j dy, s), from the
// 27-point second difference of f/g sampled at z in {s - dz, s, s + dz}, or
// the point dipole where i >= 40, j >= 40 or i^2 + j^2 >= 40^2 (the
// reference decides on the integer offsets, not on s).  In x and y the
// components keep their parity (xz odd in x, yz odd in y), so the 2-D
// spectrum is real for xx, yy, zz, xy and imaginary for xz, yz.  Output:
// 6 x nkx x (ny+1) values [c][kx][ky]: the real parts of xx,yy,zz,xy and
// the imaginary parts of xz, yz.
std::vector<double> newell_layer_spectrum(int nx, int ny, double dx, double dy, double dz, double s) {
    const int fx = std::min(nx, kFar) + 2, fy = std::min(ny, kFar) + 2, fz = 3;
    const size_t fl = (size_t)fx * fy * fz;
    std::vector<double> pv[6];
    for (auto &v : pv) v.resize(fl);
    for (int k = 0; k < fz; ++k)
        for (int j = 0; j < fy; ++j)
            for (int i = 0; i < fx; ++i) {
                const double x = (i - 1) * dx, y = (j - 1) * dy, z = (k - 1) * dz + s;
                const size_t id = (size_t)k * fy * fx + (size_t)j * fx + i;
                pv[0][id] = pot_f(x, y, z);
                pv[1][id] = pot_f(y, x, z);
                pv[2][id] = pot_f(z, y, x);
                pv[3][id] = pot_g(x, y, z);
                pv[4][id] = pot_g(x, z, y);
                pv[5][id] = pot_g(y, z, x);
            }
    const size_t n = (size_t)nx * ny;
    std::vector<double> tens(6 * n);
    const double vol = dx * dy * dz, den = 4.0 * kPi * vol;
    parallel_for(ny, [&](long lo, long hi) {
        for (long j = lo; j < hi; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t o = (size_t)j * nx + i;
                if (i >= kFar || j >= kFar || (long)i * i + j * j >= (long)kFar * kFar) {
                    const double x = i * dx, y = j * dy, z = s;
                    const double r2 = x * x + y * y + z * z, r = std::sqrt(r2);
                    const double r3 = r2 * r, r5 = r2 * r2 * r;
                    auto dia = [&](double a) { return (1.0 / r3 - 3.0 * a * a / (r3 * r2)) / (4.0 * kPi) * vol; };
                    auto off = [&](double a, double b) { return -3.0 * a * b / (4.0 * kPi * r5) * vol; };
                    tens[o] = dia(x);
                    tens[n + o] = dia(y);
                    tens[2 * n + o] = dia(z);
                    tens[3 * n + o] = off(x, y);
                    tens[4 * n + o] = off(x, z);
                    tens[5 * n + o] = off(y, z);
                } else {
                    for (int c = 0; c < 6; ++c) tens[c * n + o] = second_difference(pv[c], fx, fy, i, (int)j, 0) / den;
                }
            }
    });
    const int nkx = nx + 1, nky = ny + 1;
    const size_t red = (size_t)nkx * nky;
    std::vector<double> out(6 * red);
    const int par[6][2] = {{1, 1}, {1, 1}, {1, 1}, {-1, -1}, {-1, 1}, {1, -1}};
    for (int c = 0; c < 6; ++c) {
        std::vector<cd> a(n);
        for (size_t i = 0; i < n; ++i) a[i] = tens[c * n + i];
        std::vector<cd> b((size_t)nkx * ny);
        axis_transform(a, b, nx, ny, 1, 1, nx, 0, 1, nkx, 0, par[c][0]);
        std::vector<cd> d((size_t)nkx * nky);
        axis_transform(b, d, ny, nkx, 1, nkx, 1, 0, nkx, 1, 0, par[c][1]);
        for (int kx = 0; kx < nkx; ++kx)
            for (int ky = 0; ky < nky; ++ky) {
                const cd v = d[(size_t)ky * nkx + kx];
                out[c * red + (size_t)kx * nky + ky] = c < 4 ? v.real() : v.imag();
            }
    }
    return out;
}

// Spectrum with periodic axes (crates/fullmag-engine/src/fdm_fft.rs:140-229,
// 611-717).  On a periodic axis of n cells the reference convolves
// circularly over n with the truncated image sum
//   K(i) = sum_{m=-I..I} N(i + m n),  i in [0, n)
// (images on both sides of the base offset i >= 0, so K is not even and its
// spectrum is complex).  The engine keeps its zero-padded length 2n on that
// axis and places K there so that the padded linear convolution of the n
// data points equals the circular one: padded index p < n holds K(p),
// p > n holds K(p - n) (offset p - 2n = (p - n) - n), index n is never used.
// Open axes are as in the open tensor; a thin film (nz == 1) keeps its 2-D
// transform, with the z images summed when z is periodic.  Output, complex
// interleaved: thin film [c][kx][ky] (c = xx,yy,zz,xy), 3-D [kx][kz][ky][c]
// (c = xx,yy,zz,xy,xz,yz), kx in [0, nx], ky in [0, 2ny), kz in [0, 2nz).
bool newell_periodic_spectrum(int nx, int ny, int nz, double dx, double dy, double dz, const int per[3],
                              const int img[3], std::vector<double> &out, std::string *err) {
    const int n[3] = {nx, ny, nz};
    const bool film = nz == 1;
    int ext[3], I[3];
    double cells = 1.0;
    for (int a = 0; a < 3; ++a) {
        I[a] = per[a] ? img[a] : 0;
        ext[a] = n[a] * (I[a] + 1);
        cells *= ext[a];
    }
    if (cells > (double)(1 << 26)) {
        *err = "periodic image sum too large (cells x (images + 1) per periodic axis above 2^26)";
        return false;
    }
    const std::vector<double> oct = newell_octant(ext[0], ext[1], ext[2], dx, dy, dz);
    const size_t on = (size_t)ext[0] * ext[1] * ext[2];
    const int ncomp = film ? 4 : 6;
    const int px = 2 * nx, py = 2 * ny, pz = film ? 1 : 2 * nz;
    const int P[3] = {px, py, pz};
    // base offset of padded index p on axis a (INT_MIN: unused index)
    auto base = [&](int a, int p) -> int {
        if (a == 2 && film) return 0;
        if (p < n[a]) return p;
        if (p == n[a]) return INT_MIN;
        return per[a] ? p - n[a] : p - 2 * n[a];
    };
    // parity of the odd axes per component: xx,yy,zz even; xy odd in x,y;
    // xz odd in x,z; yz odd in y,z
    const int odd[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    const size_t PP = (size_t)px * py * pz;
    std::vector<cd> k2((size_t)ncomp * PP, cd(0.0));
    parallel_for((long)pz * py, [&](long lo, long hi) {
        for (long l = lo; l < hi; ++l) {
            const long pk = l / py;
            const int pj = (int)(l % py);
                for (int pi = 0; pi < px; ++pi) {
                    const int b[3] = {base(0, pi), base(1, pj), base(2, (int)pk)};
                    if (b[0] == INT_MIN || b[1] == INT_MIN || b[2] == INT_MIN) continue;
                    double sum[6] = {0, 0, 0, 0, 0, 0};
                    const int step_z = film ? 1 : n[2];
                    for (int mz = -I[2]; mz <= I[2]; ++mz)
                        for (int my = -I[1]; my <= I[1]; ++my)
                            for (int mx = -I[0]; mx <= I[0]; ++mx) {
                                const int o[3] = {b[0] + mx * n[0], b[1] + my * n[1], b[2] + mz * step_z};
                                const size_t id = ((size_t)std::abs(o[2]) * ext[1] + std::abs(o[1])) * ext[0] +
                                                  std::abs(o[0]);
                                for (int c = 0; c < ncomp; ++c) {
                                    double s = 1.0;
                                    for (int a = 0; a < 3; ++a)
                                        if (odd[c][a] && o[a] < 0) s = -s;
                                    sum[c] += s * oct[c * on + id];
                                }
                            }
                    const size_t q = ((size_t)pk * py + pj) * px + pi;
                    for (int c = 0; c < ncomp; ++c) k2[c * PP + q] = sum[c];
                }
        }
    });
    // complex FFT along x (kx in [0, nx] kept), y and z (all kept)
    const int nkx = nx + 1;
    std::vector<cd> sx((size_t)ncomp * nkx * py * pz);
    {
        LineFft f(px);
        parallel_for((long)ncomp * py * pz, [&](long lo, long hi) {
            std::vector<cd> a(px), t(px);
            for (long l = lo; l < hi; ++l) {
                const long c = l / ((long)py * pz), r = l % ((long)py * pz);
                for (int i = 0; i < px; ++i) a[i] = k2[c * PP + (size_t)r * px + i];
                f.run(a.data(), t.data());
                for (int k = 0; k < nkx; ++k) sx[((size_t)c * py * pz + r) * nkx + k] = a[k];
            }
        });
    }
    // layout now [c][z][y][kx]; y transform
    {
        LineFft f(py);
        parallel_for((long)ncomp * pz * nkx, [&](long lo, long hi) {
            std::vector<cd> a(py), t(py);
            for (long l = lo; l < hi; ++l) {
                const long c = l / ((long)pz * nkx), r = l % ((long)pz * nkx), z = r / nkx, kx = r % nkx;
                auto at = [&](int y) -> cd & { return sx[(((size_t)c * pz + z) * py + y) * nkx + kx]; };
                for (int y = 0; y < py; ++y) a[y] = at(y);
                f.run(a.data(), t.data());
                for (int y = 0; y < py; ++y) at(y) = a[y];
            }
        });
    }
    if (pz > 1) {
        LineFft f(pz);
        parallel_for((long)ncomp * py * nkx, [&](long lo, long hi) {
            std::vector<cd> a(pz), t(pz);
            for (long l = lo; l < hi; ++l) {
                const long c = l / ((long)py * nkx), r = l % ((long)py * nkx), y = r / nkx, kx = r % nkx;
                auto at = [&](int z) -> cd & { return sx[(((size_t)c * pz + z) * py + y) * nkx + kx]; };
                for (int z = 0; z < pz; ++z) a[z] = at(z);
                f.run(a.data(), t.data());
                for (int z = 0; z < pz; ++z) at(z) = a[z];
            }
        });
    }
    out.assign(2 * sx.size(), 0.0);
    for (int c = 0; c < ncomp; ++c)
        for (int z = 0; z < pz; ++z)
            for (int y = 0; y < py; ++y)
                for (int kx = 0; kx < nkx; ++kx) {
                    const cd v = sx[(((size_t)c * pz + z) * py + y) * nkx + kx];
                    const size_t o = film ? ((size_t)c * nkx + kx) * py + y
                                          : (((size_t)kx * pz + z) * py + y) * ncomp + c;
                    out[2 * o] = v.real();
                    out[2 * o + 1] = v.imag();
                }
    return true;
}

// Reduced spectrum from reference-format spectra (complex interleaved over the
// padded grid, thin-film 2-D or 3-D form).  Real parts are sampled at the
// kept frequencies.  Returns false (with *err) on a length mismatch.
bool reduce_given_spectra(const double *const spec[6], uint64_t len, int nx, int ny, int nz,
                          std::vector<double> &out, std::string *err) {
    const uint64_t px = 2ull * nx, py = 2ull * ny, pz3 = 2ull * nz;
    const bool film = nz == 1;
    // a film accepts the 2-D form (pz = 1) or the 3-D form; only the kz = 0
    // plane is sampled either way
    if (!(film && len == 2 * px * py) && len != 2 * px * py * pz3) {
        *err = "demag_kernel_spectrum_len mismatch: expected " + std::to_string(2 * px * py * pz3) +
               (film ? " or " + std::to_string(2 * px * py) : std::string()) + ", got " + std::to_string(len);
        return false;
    }
    const int ncomp = film ? 4 : 6;
    for (int c = 0; c < ncomp; ++c)
        if (!spec[c]) {
            *err = "demag kernel spectra must be given for all components or none";
            return false;
        }
    const int nkx = nx + 1, nky = ny + 1, nkz = film ? 1 : nz + 1;
    const size_t red = (size_t)nkx * nky * nkz;
    out.assign((size_t)ncomp * red, 0.0);
    for (int c = 0; c < ncomp; ++c)
        for (int kx = 0; kx < nkx; ++kx)
            for (int ky = 0; ky < nky; ++ky)
                for (int kz = 0; kz < nkz; ++kz) {
                    const uint64_t q = ((uint64_t)kz * py + ky) * px + kx;
                    out[c * red + ((size_t)kx * nky + ky) * nkz + kz] = spec[c][2 * q];
                }
    return true;
}

}  // namespace fm
