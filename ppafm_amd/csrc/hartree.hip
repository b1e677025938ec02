// Grids of the Hartree and full-density-based force fields on gfx950.
//
// Restated from ppafm/ocl (field.py DataGrid family, FFTCrossCorrelation,
// ForceField_LJC.calc_force_hartree / calc_force_fdbm / run_gradPotentialGrid
// and the FF.cl kernels they launch), with every value and coordinate in
// FP64 where the reference keeps FP32 grids:
//   interp_at            FF.cl:1010-1046   periodic trilinear, rotated points
//   interp_tip_at        FF.cl:1049-1093   tip centred on the origin
//   grad                 FF.cl:920-961     periodic centred differences
//   gradPotentialGrid    FF.cl:964-1007    -grad V from interpolated V
//   evalLJC_Hartree      FF.cl:860-917     tip charges in the field (the
//                                          Lennard-Jones part is the grid
//                                          kernel of ff_grid.hip)
//   power / addMult / clamp_*             FF.cl:1096-1230
//   normalizeSumReduce + sumSingleGroup   FF.cl:1112-1176
//   MultipoleTipDensity  field.py:668-705 + fieldFFT.py:99-118
//   FFTCrossCorrelation  field.py:707-843 (reikna complex64 FFTs there;
//                        hipFFT real-to-complex FP64 transforms here)
//
// Every kernel is one thread per grid point and bound by HBM: the trilinear
// gathers read 8 neighbours that mostly hit L2 (consecutive threads walk
// along z), the rest stream their inputs once.

#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "pp_common.hpp"

#pragma clang fp contract(off)

namespace pp {
namespace {

constexpr int kBlock = 256;

// A periodic grid as the trilinear interpolation sees it: origin, rows of
// (step^-1)^T (a Cartesian offset dotted with row r is index coordinate r)
// and the size.
struct Lat {
    double o[3];
    double T[9];
    int n[3];
    double inv_n[3];
};

// The grid points of the output side of a kernel: origin and step vectors.
struct Pts {
    double o[3];
    double s[9];  // rows: step vectors
    int n[3];
};

// f mod n for an integral double f (a floor): the quotient from one
// multiply, one correction step, then a clamp that keeps any input (huge,
// infinite or NaN coordinates) inside the grid.  No integer division.
__device__ __forceinline__ int wrap_index(double f, int n, double inv_n)
{
    const double q = floor(f * inv_n);
    int i = (int)(f - q * n);
    if (i >= n) i -= n;
    if (i < 0) i += n;
    return min(max(i, 0), n - 1);
}

// fract() of the OpenCL kernels: x - floor(x), kept below 1.
__device__ __forceinline__ double fract_d(double x, double &fl)
{
    fl = floor(x);
    return fmin(x - fl, 0x1.fffffffffffffp-1);
}

struct Corners {
    int i[8];
    double dx, dy, dz;
};

// linearInterpB (FF.cl:33-71): the voxel holding p and its far corner,
// indices periodic; corner order 000, 001, 010, 011, 100, 101, 110, 111.
__device__ __forceinline__ Corners corners(const Lat &L, double px, double py, double pz)
{
    const double cx = px - L.o[0], cy = py - L.o[1], cz = pz - L.o[2];
    const double u = cx * L.T[0] + cy * L.T[1] + cz * L.T[2];
    const double v = cx * L.T[3] + cy * L.T[4] + cz * L.T[5];
    const double w = cx * L.T[6] + cy * L.T[7] + cz * L.T[8];
    double fu, fv, fw;
    Corners c;
    c.dx = fract_d(u, fu);
    c.dy = fract_d(v, fv);
    c.dz = fract_d(w, fw);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const int i0 = wrap_index(fu, nx, L.inv_n[0]), j0 = wrap_index(fv, ny, L.inv_n[1]), k0 = wrap_index(fw, nz, L.inv_n[2]);
    const int i1 = i0 + 1 == nx ? 0 : i0 + 1, j1 = j0 + 1 == ny ? 0 : j0 + 1, k1 = k0 + 1 == nz ? 0 : k0 + 1;
    // grids hold fewer than 2^31 points (checked on the host)
    const int a0 = i0 * ny, a1 = i1 * ny;
    const int r00 = (a0 + j0) * nz, r01 = (a0 + j1) * nz, r10 = (a1 + j0) * nz, r11 = (a1 + j1) * nz;
    c.i[0] = r00 + k0;
    c.i[1] = r00 + k1;
    c.i[2] = r01 + k0;
    c.i[3] = r01 + k1;
    c.i[4] = r10 + k0;
    c.i[5] = r10 + k1;
    c.i[6] = r11 + k0;
    c.i[7] = r11 + k1;
    return c;
}

// The weights in the reference's order of terms.
__device__ __forceinline__ void weights(const Corners &c, double w[8])
{
    const double ex = 1 - c.dx, ey = 1 - c.dy, ez = 1 - c.dz;
    w[0] = ex * ey * ez;
    w[1] = ex * ey * c.dz;
    w[2] = ex * c.dy * ez;
    w[3] = ex * c.dy * c.dz;
    w[4] = c.dx * ey * ez;
    w[5] = c.dx * ey * c.dz;
    w[6] = c.dx * c.dy * ez;
    w[7] = c.dx * c.dy * c.dz;
}

__device__ __forceinline__ double trilinear(const double *__restrict__ buf, const Lat &L, double px, double py,
                                            double pz)
{
    const Corners c = corners(L, px, py, pz);
    double w[8];
    weights(c, w);
    double s = w[0] * buf[c.i[0]];
    for (int q = 1; q < 8; q++) s += w[q] * buf[c.i[q]];
    return s;
}

__device__ __forceinline__ double4 trilinear4(const double4 *__restrict__ buf, const Lat &L, double px, double py,
                                              double pz)
{
    const Corners c = corners(L, px, py, pz);
    double w[8];
    weights(c, w);
    double4 s = make_double4(0, 0, 0, 0);
    for (int q = 0; q < 8; q++) {
        const double4 v = buf[c.i[q]];
        s.x += w[q] * v.x;
        s.y += w[q] * v.y;
        s.z += w[q] * v.z;
        s.w += w[q] * v.w;
    }
    return s;
}

// Grid indices of a C-order point number (32-bit unsigned arithmetic:
// grids hold fewer than 2^31 points).
__device__ __forceinline__ void split_c(long ind, const int n[3], int &i, int &j, int &k)
{
    const unsigned u = (unsigned)ind;
    const unsigned r = u / (unsigned)n[2];
    k = (int)(u - r * (unsigned)n[2]);
    const unsigned q = r / (unsigned)n[1];
    j = (int)(r - q * (unsigned)n[1]);
    i = (int)q;
}

// origin + i a' + j b' + k c', summed in that order.
__device__ __forceinline__ void point(const Pts &P, int i, int j, int k, double &x, double &y, double &z)
{
    x = P.o[0] + i * P.s[0] + j * P.s[3] + k * P.s[6];
    y = P.o[1] + i * P.s[1] + j * P.s[4] + k * P.s[7];
    z = P.o[2] + i * P.s[2] + j * P.s[5] + k * P.s[8];
}

__device__ __forceinline__ long npts(const int n[3]) { return (long)n[0] * n[1] * n[2]; }

// ----------------------------------------------------------------- kernels

struct Rot {
    double r[9];  // row-major
    double c[3];  // centre
};

// interp_at (FF.cl:1010-1046).
__global__ __launch_bounds__(kBlock) void interp_at_kernel(const double *__restrict__ in, Lat L, Pts P, Rot R,
                                                           double *__restrict__ out)
{
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= npts(P.n)) return;
    int i, j, k;
    split_c(ind, P.n, i, j, k);
    double x, y, z;
    point(P, i, j, k, x, y, z);
    const double dx = x - R.c[0], dy = y - R.c[1], dz = z - R.c[2];
    const double qx = R.c[0] + (R.r[0] * dx + R.r[1] * dy + R.r[2] * dz);
    const double qy = R.c[1] + (R.r[3] * dx + R.r[4] * dy + R.r[5] * dz);
    const double qz = R.c[2] + (R.r[6] * dx + R.r[7] * dy + R.r[8] * dz);
    out[ind] = trilinear(in, L, qx, qy, qz);
}

// interp_tip_at (FF.cl:1049-1093); L.o is zero, inv = rows of inv(A).
struct Inv3 {
    double m[9];
};

__global__ __launch_bounds__(kBlock) void interp_tip_at_kernel(const double *__restrict__ in, Lat L, Pts P, Inv3 V,
                                                               double *__restrict__ out)
{
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= npts(P.n)) return;
    int i, j, k;
    split_c(ind, P.n, i, j, k);
    if (i > P.n[0] / 2) i -= P.n[0];
    if (j > P.n[1] / 2) j -= P.n[1];
    if (k > P.n[2] / 2) k -= P.n[2];
    const double x = i * P.s[0] + j * P.s[3] + k * P.s[6];
    const double y = i * P.s[1] + j * P.s[4] + k * P.s[7];
    const double z = i * P.s[2] + j * P.s[5] + k * P.s[8];
    const double fa = V.m[0] * x + V.m[1] * y + V.m[2] * z;
    const double fb = V.m[3] * x + V.m[4] * y + V.m[5] * z;
    const double fc = V.m[6] * x + V.m[7] * y + V.m[8] * z;
    if (fabs(fa) > 0.5 || fabs(fb) > 0.5 || fabs(fc) > 0.5)
        out[ind] = 0.0;
    else
        out[ind] = trilinear(in, L, x, y, z);
}

struct GradArgs {
    int n[3];
    double step[3];
    double scale[4];
    int order_c, accumulate;
};

// grad (FF.cl:920-961).
__global__ __launch_bounds__(kBlock) void grad_kernel(const double *__restrict__ in, GradArgs A,
                                                      double4 *__restrict__ out)
{
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= npts(A.n)) return;
    int i, j, k;
    split_c(ind, A.n, i, j, k);
    const long nz = A.n[2], nyz = (long)A.n[1] * nz;
    const long ip = i > 0 ? ind - nyz : ind + (A.n[0] - 1) * nyz;
    const long jp = j > 0 ? ind - nz : ind + (A.n[1] - 1) * nz;
    const long kp = k > 0 ? ind - 1 : ind + (nz - 1);
    const long in_ = i < A.n[0] - 1 ? ind + nyz : ind - (A.n[0] - 1) * nyz;
    const long jn = j < A.n[1] - 1 ? ind + nz : ind - (A.n[1] - 1) * nz;
    const long kn = k < nz - 1 ? ind + 1 : ind - (nz - 1);
    double4 f;
    f.x = A.scale[0] * (0.5 * (in[in_] - in[ip]) / A.step[0]);
    f.y = A.scale[1] * (0.5 * (in[jn] - in[jp]) / A.step[1]);
    f.z = A.scale[2] * (0.5 * (in[kn] - in[kp]) / A.step[2]);
    f.w = A.scale[3] * in[ind];
    const long o = A.order_c ? ind : i + (long)A.n[0] * j + (long)A.n[0] * A.n[1] * k;
    if (A.accumulate) {
        double4 g = out[o];
        g.x += f.x;
        g.y += f.y;
        g.z += f.z;
        g.w += f.w;
        out[o] = g;
    } else {
        out[o] = f;
    }
}

// gradPotentialGrid (FF.cl:964-1007).
__global__ __launch_bounds__(kBlock) void grad_potential_kernel(const double *__restrict__ pot, Lat L, Pts P,
                                                                double h, double4 *__restrict__ field)
{
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= npts(P.n)) return;
    int i, j, k;
    split_c(ind, P.n, i, j, k);
    double x, y, z;
    point(P, i, j, k, x, y, z);
    const double v = trilinear(pot, L, x, y, z);
    const double vip = trilinear(pot, L, x + (-h), y, z);
    const double vjp = trilinear(pot, L, x, y + (-h), z);
    const double vkp = trilinear(pot, L, x, y, z + (-h));
    const double vin = trilinear(pot, L, x + h, y, z);
    const double vjn = trilinear(pot, L, x, y + h, z);
    const double vkn = trilinear(pot, L, x, y, z + h);
    field[ind] = make_double4(0.5 * (vip - vin) / h, 0.5 * (vjp - vjn) / h, 0.5 * (vkp - vkn) / h, v);
}

struct Charges {
    double q[4], dz[4];
    int nq;
};

// The tip-charge terms of evalLJC_Hartree (FF.cl:906-911), added in order.
__global__ __launch_bounds__(kBlock) void tip_charges_kernel(const double4 *__restrict__ field, Lat L, Pts P,
                                                             Charges C, double4 *__restrict__ grid)
{
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= npts(P.n)) return;
    int i, j, k;
    split_c(ind, P.n, i, j, k);
    double x, y, z;
    point(P, i, j, k, x, y, z);
    double4 g = grid[ind];
    for (int q = 0; q < C.nq; q++) {
        const double4 e = trilinear4(field, L, x, y, z + C.dz[q]);
        g.x += C.q[q] * e.x;
        g.y += C.q[q] * e.y;
        g.z += C.q[q] * e.z;
        g.w += C.q[q] * e.w;
    }
    grid[ind] = g;
}

// Sums (total, positive part) in two fixed-shape passes.
constexpr int kSumBlocks = 1024;

__device__ __forceinline__ void block_sum2(double &a, double &b)
{
    __shared__ double sa[kBlock], sb[kBlock];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
}

__global__ __launch_bounds__(kBlock) void sums_partial_kernel(const double *__restrict__ in, long n,
                                                              double2 *__restrict__ part)
{
    double tot = 0, pos = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = in[i];
        tot += v;
        pos += v > 0 ? v : 0.0;
    }
    block_sum2(tot, pos);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(tot, pos);
}

__global__ __launch_bounds__(kBlock) void sums_final_kernel(double2 *__restrict__ part, int nparts)
{
    double tot = 0, pos = 0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        tot += part[i].x;
        pos += part[i].y;
    }
    block_sum2(tot, pos);
    if (threadIdx.x == 0) part[0] = make_double2(tot, pos);
}

__global__ __launch_bounds__(kBlock) void power_kernel(const double *__restrict__ in, double *__restrict__ out, long n,
                                                       double p, double scale)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = scale * pow(fmax(0.0, in[i]), p);
}

__global__ __launch_bounds__(kBlock) void add_mult_kernel(const double *__restrict__ a, const double *__restrict__ b,
                                                          double *__restrict__ out, long n, double A)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = a[i] + A * b[i];
}

__global__ __launch_bounds__(kBlock) void clamp_kernel(const double *__restrict__ in, double *__restrict__ out, long n,
                                                       double lo, double hi, int soft, double width)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = in[i];
    if (!soft) {
        v = fmax(lo, fmin(hi, v));
    } else {
        if (hi < INFINITY) v = (v - hi) / (1 + exp((v - hi) / width)) + hi;
        if (lo > -INFINITY) v = (v - lo) / (1 + exp((v - lo) / -width)) + lo;
    }
    out[i] = v;
}

constexpr int kMaxKinds = 16;

struct TipArgs {
    int n[3];
    double mat[9];
    double sigma, two_s2, ca, sa;
    double norm[4];
    int nkinds;
    int kinds[kMaxKinds];
    double coefs[kMaxKinds];
};

// getSphericalHarmonic (fieldFFT.py:58-96) after rotZX by the tilt.
__device__ __forceinline__ double harmonic(int kind, double X, double Y, double Z, double ca, double sa)
{
    const double Zr = ca * Z - sa * X;
    const double Xr = sa * Z + ca * X;
    switch (kind) {
    case 0: return 1.0;
    case 1: return Xr;
    case 2: return Y;
    case 3: return Zr;
    case 4: return 0.25 * (2 * (Zr * Zr) - Xr * Xr - Y * Y);
    case 5: return 0.25 * (2 * (Xr * Xr) - Y * Y - Zr * Zr);
    case 6: return 0.25 * (2 * (Y * Y) - Xr * Xr - Zr * Zr);
    case 7: return Xr * Y;
    case 8: return Xr * Zr;
    case 9: return Y * Zr;
    default: return 0.0;
    }
}

// pass 0: out = exp(-|r|^2 / (2 sigma^2)); pass 1: out = out / (sum *
// norm...) times the multipole combination (getProbeDensity,
// fieldFFT.py:99-118).
__global__ __launch_bounds__(kBlock) void tip_density_kernel(TipArgs A, const double *__restrict__ cx,
                                                             const double *__restrict__ cy,
                                                             const double *__restrict__ cz,
                                                             const double2 *__restrict__ sums, int pass,
                                                             double *__restrict__ out)
{
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= npts(A.n)) return;
    int i, j, k;
    split_c(ind, A.n, i, j, k);
    const double X = cx[i], Y = cy[j], Z = cz[k];
    const double *m = A.mat;
    const double rx = X * m[0] + Y * m[1] + Z * m[2];
    const double ry = X * m[3] + Y * m[4] + Z * m[5];
    const double rz = X * m[6] + Y * m[7] + Z * m[8];
    if (pass == 0) {
        const double rquad = rx * rx + ry * ry + rz * rz;
        out[ind] = exp(-rquad / A.two_s2);
        return;
    }
    const double renorm = sums[0].x * A.norm[0] * A.norm[1] * A.norm[2] * A.norm[3];
    const double radial = out[ind] / renorm;
    if (A.nkinds < 0) {
        out[ind] = radial;
        return;
    }
    const double hx = rx / A.sigma, hy = ry / A.sigma, hz = rz / A.sigma;
    double rho = 0.0;
    for (int q = 0; q < A.nkinds; q++) rho += radial * A.coefs[q] * harmonic(A.kinds[q], hx, hy, hz, A.ca, A.sa);
    out[ind] = rho;
}

// conj(rho_hat) * in_hat * s over the half spectrum (the complex product
// of FFTCrossCorrelation.conj_mult, field.py:760-773).
__global__ __launch_bounds__(kBlock) void conj_mult_kernel(const double2 *__restrict__ rho,
                                                           double2 *__restrict__ x, long n, double s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 a = rho[i], b = x[i];
    x[i] = make_double2((a.x * b.x + a.y * b.y) * s, (a.x * b.y - a.y * b.x) * s);
}

__global__ __launch_bounds__(kBlock) void copy_kernel(const double *__restrict__ in, double *__restrict__ out, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// ----------------------------------------------------------------- helpers

// Positive sizes and fewer than 2^31 points (the kernels index in 32 bits).
bool valid_n(const int n[3])
{
    return n && n[0] > 0 && n[1] > 0 && n[2] > 0 && (long long)n[0] * n[1] * n[2] < (1LL << 31);
}

unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

bool inv3(const double a[9], double r[9])
{
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(det != 0) || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    r[0] = c00 * id;
    r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    r[3] = c01 * id;
    r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    r[6] = c02 * id;
    r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return true;
}

// Step vectors (rows) of a lattice lvec[12] sampled n times.
void steps(const double lvec[12], const int n[3], double s[9])
{
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) s[3 * r + c] = lvec[3 * (r + 1) + c] / n[r];
}

// The interpolation view of a grid: T = (steps^-1)^T as rows.
int make_lat(const double lvec[12], const int n[3], bool zero_origin, Lat &L)
{
    double s[9], si[9];
    steps(lvec, n, s);
    PP_REQUIRE(inv3(s, si), "grid lattice vectors are singular");
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) L.T[3 * r + c] = si[3 * c + r];
    for (int c = 0; c < 3; c++) {
        L.o[c] = zero_origin ? 0.0 : lvec[c];
        L.n[c] = n[c];
        L.inv_n[c] = 1.0 / n[c];
    }
    return 0;
}

void make_pts(const double lvec[12], const int n[3], Pts &P)
{
    steps(lvec, n, P.s);
    for (int c = 0; c < 3; c++) {
        P.o[c] = lvec[c];
        P.n[c] = n[c];
    }
}

long grid_points(const int n[3]) { return (long)n[0] * n[1] * n[2]; }

#define PP_FFT(expr)                                                                           \
    do {                                                                                       \
        hipfftResult _r = (expr);                                                              \
        if (_r != HIPFFT_SUCCESS) {                                                            \
            ::pp::set_error(std::string(#expr) + ": hipFFT error " + std::to_string((int)_r)); \
            return -1;                                                                         \
        }                                                                                      \
    } while (0)

}  // namespace
}  // namespace pp

using namespace pp;

struct ppafm_fft_corr {
    int n[3];
    long npts, nhat;
    hipfftHandle fwd = 0, inv = 0;
    bool have_fwd = false, have_inv = false;
    double *real = nullptr;    // npts
    double2 *hat = nullptr;    // nhat
    double2 *rho_hat = nullptr;
    double cell_vol = 0;
    bool have_rho = false;
};

extern "C" {

int ppafm_grid_interp_at(const double in_lvec[12], const int in_n[3], const double *in_dev, const double out_lvec[12],
                         const int out_n[3], double *out_dev, const double rot[9], const double rot_center[3],
                         void *stream)
{
    PP_REQUIRE(in_lvec && out_lvec && in_dev && out_dev && rot && rot_center, "null argument");
    PP_REQUIRE((const void *)in_dev != (const void *)out_dev, "input and output grids must be different arrays");
    PP_REQUIRE(valid_n(in_n) && valid_n(out_n), "grid sizes must be positive with fewer than 2^31 points");
    Lat L;
    Pts P;
    Rot R;
    if (int e = make_lat(in_lvec, in_n, false, L)) return e;
    make_pts(out_lvec, out_n, P);
    for (int i = 0; i < 9; i++) R.r[i] = rot[i];
    for (int i = 0; i < 3; i++) R.c[i] = rot_center[i];
    const long n = grid_points(out_n);
    hipLaunchKernelGGL(interp_at_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, as_stream(stream), in_dev, L, P, R,
                       out_dev);
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_grid_interp_tip_at(const double in_lvec[12], const int in_n[3], const double *in_dev,
                             const double out_lvec[12], const int out_n[3], double *out_dev, void *stream)
{
    PP_REQUIRE(in_lvec && out_lvec && in_dev && out_dev, "null argument");
    PP_REQUIRE((const void *)in_dev != (const void *)out_dev, "input and output grids must be different arrays");
    PP_REQUIRE(valid_n(in_n) && valid_n(out_n), "grid sizes must be positive with fewer than 2^31 points");
    Lat L;
    Pts P;
    Inv3 V;
    if (int e = make_lat(in_lvec, in_n, true, L)) return e;
    make_pts(out_lvec, out_n, P);
    PP_REQUIRE(inv3(in_lvec + 3, V.m), "tip grid lattice vectors are singular");
    const long n = grid_points(out_n);
    hipLaunchKernelGGL(interp_tip_at_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, as_stream(stream), in_dev, L, P, V,
                       out_dev);
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_grid_grad(const int n[3], const double step[3], const double *in_dev, const double scale[4], int order_c,
                    int accumulate, double *out_dev, void *stream)
{
    PP_REQUIRE(step && in_dev && scale && out_dev, "null argument");
    PP_REQUIRE((const void *)in_dev != (const void *)out_dev, "input and output grids must be different arrays");
    PP_REQUIRE(valid_n(n), "grid sizes must be positive with fewer than 2^31 points");
    GradArgs A;
    for (int i = 0; i < 3; i++) {
        A.n[i] = n[i];
        A.step[i] = step[i];
    }
    for (int i = 0; i < 4; i++) A.scale[i] = scale[i];
    A.order_c = order_c != 0;
    A.accumulate = accumulate != 0;
    const long np = grid_points(n);
    hipLaunchKernelGGL(grad_kernel, dim3(blocks_for(np)), dim3(kBlock), 0, as_stream(stream), in_dev, A,
                       (double4 *)out_dev);
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_grid_grad_potential(const double pot_lvec[12], const int pot_n[3], const double *pot_dev,
                              const double lvec[12], const int n[3], double h, double *field_dev, void *stream)
{
    PP_REQUIRE(pot_lvec && pot_dev && lvec && field_dev, "null argument");
    PP_REQUIRE((const void *)pot_dev != (const void *)field_dev, "input and output grids must be different arrays");
    PP_REQUIRE(valid_n(pot_n) && valid_n(n), "grid sizes must be positive with fewer than 2^31 points");
    PP_REQUIRE(h > 0, "finite-difference step must be positive");
    Lat L;
    Pts P;
    if (int e = make_lat(pot_lvec, pot_n, false, L)) return e;
    make_pts(lvec, n, P);
    const long np = grid_points(n);
    hipLaunchKernelGGL(grad_potential_kernel, dim3(blocks_for(np)), dim3(kBlock), 0, as_stream(stream), pot_dev, L, P,
                       h, (double4 *)field_dev);
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_grid_add_tip_charges(const double lvec[12], const int n[3], const double *field_dev, const double Qs[4],
                               const double QZs[4], double *grid4_dev, void *stream)
{
    PP_REQUIRE(lvec && field_dev && Qs && QZs && grid4_dev, "null argument");
    PP_REQUIRE(valid_n(n), "grid sizes must be positive with fewer than 2^31 points");
    PP_REQUIRE(field_dev != grid4_dev, "the field and the output grid must be different arrays");
    Lat L;
    Pts P;
    if (int e = make_lat(lvec, n, false, L)) return e;
    make_pts(lvec, n, P);
    Charges C;
    C.nq = 0;
    for (int q = 0; q < 4; q++)
        if (Qs[q] != 0) {
            C.q[C.nq] = Qs[q];
            C.dz[C.nq] = QZs[q];
            C.nq++;
        }
    if (C.nq == 0) return 0;
    const long np = grid_points(n);
    hipLaunchKernelGGL(tip_charges_kernel, dim3(blocks_for(np)), dim3(kBlock), 0, as_stream(stream),
                       (const double4 *)field_dev, L, P, C, (double4 *)grid4_dev);
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_grid_sums(long long n, const double *in_dev, double sums[2], void *stream)
{
    PP_REQUIRE(sums && (in_dev || n == 0), "null argument");
    PP_REQUIRE(n >= 0, "negative size");
    sums[0] = sums[1] = 0;
    if (n == 0) return 0;
    const hipStream_t st = as_stream(stream);
    const int nb = (int)std::min<long long>(kSumBlocks, (n + kBlock - 1) / kBlock);
    double2 *part = nullptr;
    PP_TRY_HIP(hipMallocAsync((void **)&part, sizeof(double2) * nb, st));
    hipLaunchKernelGGL(sums_partial_kernel, dim3(nb), dim3(kBlock), 0, st, in_dev, (long)n, part);
    hipLaunchKernelGGL(sums_final_kernel, dim3(1), dim3(kBlock), 0, st, part, nb);
    double2 h;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, part, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFreeAsync(part, st);
    PP_TRY_HIP(e);
    sums[0] = h.x;
    sums[1] = h.y;
    return 0;
}

int ppafm_grid_power(long long n, const double *in_dev, double *out_dev, double p, double scale, void *stream)
{
    PP_REQUIRE((in_dev && out_dev) || n == 0, "null argument");
    if (n <= 0) return 0;
    hipLaunchKernelGGL(power_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, as_stream(stream), in_dev, out_dev, (long)n,
                       p, scale);
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_grid_add_mult(long long n, const double *in1_dev, const double *in2_dev, double A, double *out_dev,
                        void *stream)
{
    PP_REQUIRE((in1_dev && in2_dev && out_dev) || n == 0, "null argument");
    if (n <= 0) return 0;
    hipLaunchKernelGGL(add_mult_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, as_stream(stream), in1_dev, in2_dev,
                       out_dev, (long)n, A);
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_grid_clamp(long long n, const double *in_dev, double *out_dev, double lo, double hi, int soft, double width,
                     void *stream)
{
    PP_REQUIRE((in_dev && out_dev) || n == 0, "null argument");
    if (n <= 0) return 0;
    hipLaunchKernelGGL(clamp_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, as_stream(stream), in_dev, out_dev, (long)n,
                       lo, hi, soft != 0, width);
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_tip_density_multipole(const int n[3], const double *cx, const double *cy, const double *cz,
                                const double mat[9], double sigma, double tilt, int nkinds, const int *kinds,
                                const double *coefs, const double norm[4], double *out_dev, void *stream)
{
    PP_REQUIRE(cx && cy && cz && mat && norm && out_dev, "null argument");
    PP_REQUIRE(valid_n(n), "grid sizes must be positive with fewer than 2^31 points");
    PP_REQUIRE(nkinds <= kMaxKinds, "at most 16 multipole terms");
    PP_REQUIRE(nkinds <= 0 || (kinds && coefs), "null multipole arrays");
    PP_REQUIRE(sigma > 0, "sigma must be positive");
    TipArgs A;
    for (int i = 0; i < 3; i++) A.n[i] = n[i];
    for (int i = 0; i < 9; i++) A.mat[i] = mat[i];
    A.sigma = sigma;
    A.two_s2 = 2 * (sigma * sigma);
    A.ca = std::cos(tilt);
    A.sa = std::sin(tilt);
    for (int i = 0; i < 4; i++) A.norm[i] = norm[i];
    A.nkinds = nkinds;
    for (int q = 0; q < std::max(nkinds, 0); q++) {
        A.kinds[q] = kinds[q];
        A.coefs[q] = coefs[q];
    }
    const hipStream_t st = as_stream(stream);
    const long np = grid_points(n);
    const int nb = (int)std::min<long>(kSumBlocks, (np + kBlock - 1) / kBlock);
    const size_t ncoord = (size_t)n[0] + n[1] + n[2];
    double *coord = nullptr;
    double2 *part = nullptr;
    PP_TRY_HIP(hipMallocAsync((void **)&coord, ncoord * sizeof(double), st));
    hipError_t e = hipMallocAsync((void **)&part, sizeof(double2) * nb, st);
    if (e == hipSuccess) e = hipMemcpyAsync(coord, cx, n[0] * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(coord + n[0], cy, n[1] * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(coord + n[0] + n[1], cz, n[2] * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const double *dx = coord, *dy = coord + n[0], *dz = coord + n[0] + n[1];
        hipLaunchKernelGGL(tip_density_kernel, dim3(blocks_for(np)), dim3(kBlock), 0, st, A, dx, dy, dz,
                           (const double2 *)nullptr, 0, out_dev);
        hipLaunchKernelGGL(sums_partial_kernel, dim3(nb), dim3(kBlock), 0, st, (const double *)out_dev, np, part);
        hipLaunchKernelGGL(sums_final_kernel, dim3(1), dim3(kBlock), 0, st, part, nb);
        hipLaunchKernelGGL(tip_density_kernel, dim3(blocks_for(np)), dim3(kBlock), 0, st, A, dx, dy, dz,
                           (const double2 *)part, 1, out_dev);
        e = hipGetLastError();
    }
    // the host coordinate arrays must outlive the copies
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (part) hipFreeAsync(part, st);
    hipFreeAsync(coord, st);
    PP_TRY_HIP(e);
    return 0;
}

int ppafm_fft_corr_create(const int n[3], ppafm_fft_corr **out)
{
    PP_REQUIRE(out, "null argument");
    *out = nullptr;
    PP_REQUIRE(valid_n(n), "grid sizes must be positive with fewer than 2^31 points");
    auto *c = new ppafm_fft_corr();
    for (int i = 0; i < 3; i++) c->n[i] = n[i];
    c->npts = grid_points(n);
    c->nhat = (long)n[0] * n[1] * (n[2] / 2 + 1);
    auto fail = [&](int code) {
        ppafm_fft_corr_destroy(c);
        return code;
    };
    hipfftResult r = hipfftPlan3d(&c->fwd, n[0], n[1], n[2], HIPFFT_D2Z);
    if (r != HIPFFT_SUCCESS) {
        set_error("hipfftPlan3d (D2Z) failed: " + std::to_string((int)r));
        return fail(-1);
    }
    c->have_fwd = true;
    r = hipfftPlan3d(&c->inv, n[0], n[1], n[2], HIPFFT_Z2D);
    if (r != HIPFFT_SUCCESS) {
        set_error("hipfftPlan3d (Z2D) failed: " + std::to_string((int)r));
        return fail(-1);
    }
    c->have_inv = true;
    hipError_t e = hipMalloc((void **)&c->real, sizeof(double) * c->npts);
    if (e == hipSuccess) e = hipMalloc((void **)&c->hat, sizeof(double2) * c->nhat);
    if (e == hipSuccess) e = hipMalloc((void **)&c->rho_hat, sizeof(double2) * c->nhat);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return fail(-1);
    }
    *out = c;
    return 0;
}

int ppafm_fft_corr_set_rho(ppafm_fft_corr *c, const double *rho_dev, double cell_vol, void *stream)
{
    PP_REQUIRE(c && rho_dev, "null argument");
    const hipStream_t st = as_stream(stream);
    // the transform may use its input as scratch: run it on a copy
    hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(c->npts)), dim3(kBlock), 0, st, rho_dev, c->real, c->npts);
    PP_TRY_HIP(hipGetLastError());
    PP_FFT(hipfftSetStream(c->fwd, st));
    PP_FFT(hipfftExecD2Z(c->fwd, (hipfftDoubleReal *)c->real, (hipfftDoubleComplex *)c->rho_hat));
    c->cell_vol = cell_vol;
    c->have_rho = true;
    return 0;
}

int ppafm_fft_corr_correlate(ppafm_fft_corr *c, const double *in_dev, double scale, double *E_dev, void *stream)
{
    PP_REQUIRE(c && in_dev && E_dev, "null argument");
    PP_REQUIRE(c->have_rho, "no tip density set (ppafm_fft_corr_set_rho)");
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(c->npts)), dim3(kBlock), 0, st, in_dev, c->real, c->npts);
    PP_TRY_HIP(hipGetLastError());
    PP_FFT(hipfftSetStream(c->fwd, st));
    PP_FFT(hipfftExecD2Z(c->fwd, (hipfftDoubleReal *)c->real, (hipfftDoubleComplex *)c->hat));
    const double s = scale * c->cell_vol / (double)c->npts;
    hipLaunchKernelGGL(conj_mult_kernel, dim3(blocks_for(c->nhat)), dim3(kBlock), 0, st, c->rho_hat, c->hat, c->nhat,
                       s);
    PP_TRY_HIP(hipGetLastError());
    PP_FFT(hipfftSetStream(c->inv, st));
    PP_FFT(hipfftExecZ2D(c->inv, (hipfftDoubleComplex *)c->hat, (hipfftDoubleReal *)E_dev));
    return 0;
}

int ppafm_fft_corr_destroy(ppafm_fft_corr *c)
{
    if (!c) return 0;
    if (c->have_fwd) hipfftDestroy(c->fwd);
    if (c->have_inv) hipfftDestroy(c->inv);
    if (c->real) (void)hipFree(c->real);
    if (c->hat) (void)hipFree(c->hat);
    if (c->rho_hat) (void)hipFree(c->rho_hat);
    delete c;
    return 0;
}

int ppafm_grid_add_vdw(const ppafm_grid *grid, int damp_method, int natoms, const double *pos_dev,
                       const double *coefs_dev, double *grid4_dev, void *stream)
{
    PP_REQUIRE(grid && grid4_dev, "null argument");
    PP_REQUIRE(damp_method >= -1 && damp_method <= 4, "damp_method must be -1..4");
    PP_REQUIRE(natoms >= 0 && (natoms == 0 || (pos_dev && coefs_dev)), "null atom arrays");
    PP_REQUIRE(valid_n(grid->n), "grid sizes must be positive with fewer than 2^31 points");
    // ADamp_* of FF.cl:17-21 (float literals)
    static const int models[6] = {kModelVdWNoDampOcl, kModelVdW, kModelVdWR2, kModelVdWR4, kModelVdWInvR4Ocl,
                                  kModelVdWInvR8Ocl};
    static const double adamp[6] = {0.0, (double)180.0f, (double)0.5f, (double)0.5f, (double)0.03f, (double)0.01f};
    if (natoms == 0) return 0;
    const GridGeom g = make_geom(grid->n, grid->cell);
    const int k = damp_method + 1;
    int e = launch_ff_models_strided(g, models[k], adamp[k], natoms, pos_dev, coefs_dev, 2, grid4_dev, 4,
                                     grid4_dev + 3, 4, as_stream(stream));
    if (e) return e;
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_grid_add_d3(const ppafm_grid *grid, int natoms, const double *pos_dev, const double *coefs_dev,
                      double *grid4_dev, void *stream)
{
    PP_REQUIRE(grid && grid4_dev, "null argument");
    PP_REQUIRE(natoms >= 0 && (natoms == 0 || (pos_dev && coefs_dev)), "null atom arrays");
    PP_REQUIRE(valid_n(grid->n), "grid sizes must be positive with fewer than 2^31 points");
    if (natoms == 0) return 0;
    const GridGeom g = make_geom(grid->n, grid->cell);
    int e = launch_ff_models_strided(g, kModelD3, 0.0, natoms, pos_dev, coefs_dev, 4, grid4_dev, 4, grid4_dev + 3, 4,
                                     as_stream(stream));
    if (e) return e;
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
