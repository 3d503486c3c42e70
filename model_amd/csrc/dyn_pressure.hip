// Anelastic pressure Poisson solve for gfx950 (serial periodic branch of solve_pressure,
// pressure.cpp:142-224, with poisson_fft3d 805-893 and divergence3d_periodic 1390-1424).
//
// Pipeline over the active levels kk = k-1, k in [1, NZ-1):
//   P1  divergence (fused prologue) + forward DFT along i     fields -> spec[kx][j][kk]
//   P2  forward DFT along j                                  spec -> spec (in place)
//   P3  1/(ni*nj) scaling + tridiagonal solve in k, FP64      spec -> spec (in place)
//   P4  backward DFT along j                                 spec -> spec
//   P5  backward DFT along i + real part -> PI with the k=0/NZ-1 copies and periodic halos fused
// The DFTs are Stockham auto-sort passes over an LDS tile [n][kc] (radix 4/2/3/5), batched over a
// chunk of kc consecutive levels so every global access is a contiguous kc-run of a k column.
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "dyn_internal.hpp"

namespace dyn {
namespace {


constexpr int kThreads = 256;
constexpr int kThreadsX = 512;  // the real-transform passes along i (more loads in flight per CU)
constexpr size_t kLdsBudget = 64 * 1024;

// exp(sign * 2 pi i * e / n) from the forward table (cos, -sin)
template <typename T>
__device__ __forceinline__ void twiddle(const double *tw, int e, int sign, T &wr, T &wi) {
    wr = T(tw[2 * e]);
    wi = sign < 0 ? T(tw[2 * e + 1]) : -T(tw[2 * e + 1]);
}

template <typename T>
__device__ __forceinline__ void cmul(T &ar, T &ai, T br, T bi) {
    T r = ar * br - ai * bi;
    T i = ar * bi + ai * br;
    ar = r;
    ai = i;
}

// Two or four consecutive elements of a spectrum in one vector access.  A complex value starts at an even
// element, so a pair is aligned to its size; four elements (the two complex values of a level pair
// starting at an even level) are aligned to theirs when the level count is even.
template <typename T>
struct Vec2T;
template <>
struct Vec2T<float> {
    using type = float2;
};
template <>
struct Vec2T<double> {
    using type = double2;
};
template <typename T>
__device__ __forceinline__ void ld2(const T *p, T &a, T &b) {
    typename Vec2T<T>::type v;
    __builtin_memcpy(&v, __builtin_assume_aligned(p, sizeof(v)), sizeof(v));
    a = v.x;
    b = v.y;
}
template <typename T>
__device__ __forceinline__ void st2(T *p, T a, T b) {
    typename Vec2T<T>::type v;
    v.x = a;
    v.y = b;
    __builtin_memcpy(__builtin_assume_aligned(p, sizeof(v)), &v, sizeof(v));
}
template <typename T>
__device__ __forceinline__ void ld4(const T *p, T (&x)[4]) {
    if constexpr (sizeof(T) == 4) {
        float4 v;
        __builtin_memcpy(&v, __builtin_assume_aligned(p, 16), 16);
        x[0] = v.x;
        x[1] = v.y;
        x[2] = v.z;
        x[3] = v.w;
    } else {
        ld2(p, x[0], x[1]);
        ld2(p + 2, x[2], x[3]);
    }
}
template <typename T>
__device__ __forceinline__ void st4(T *p, T a, T b, T c, T d) {
    if constexpr (sizeof(T) == 4) {
        float4 v;
        v.x = a;
        v.y = b;
        v.z = c;
        v.w = d;
        __builtin_memcpy(__builtin_assume_aligned(p, 16), &v, 16);
    } else {
        st2(p, a, b);
        st2(p + 2, c, d);
    }
}

// in-register DFT of length R with exponent sign `sign`
template <typename T>
__device__ __forceinline__ void butterfly(int R, T *xr, T *xi, int sign) {
    const T s = T(sign);
    if (R == 2) {
        T ar = xr[0] + xr[1], ai = xi[0] + xi[1];
        xr[1] = xr[0] - xr[1];
        xi[1] = xi[0] - xi[1];
        xr[0] = ar;
        xi[0] = ai;
    } else if (R == 4) {
        T t0r = xr[0] + xr[2], t0i = xi[0] + xi[2];
        T t1r = xr[0] - xr[2], t1i = xi[0] - xi[2];
        T t2r = xr[1] + xr[3], t2i = xi[1] + xi[3];
        T t3r = xr[1] - xr[3], t3i = xi[1] - xi[3];
        // i*s*(a1 - a3) = (-s*t3i, s*t3r)
        xr[0] = t0r + t2r;
        xi[0] = t0i + t2i;
        xr[2] = t0r - t2r;
        xi[2] = t0i - t2i;
        xr[1] = t1r - s * t3i;
        xi[1] = t1i + s * t3r;
        xr[3] = t1r + s * t3i;
        xi[3] = t1i - s * t3r;
    } else if (R == 3) {
        const T m = s * T(0.86602540378443864676);
        T t1r = xr[1] + xr[2], t1i = xi[1] + xi[2];
        T dr = xr[1] - xr[2], di = xi[1] - xi[2];
        T t2r = xr[0] - T(0.5) * t1r, t2i = xi[0] - T(0.5) * t1i;
        xr[0] = xr[0] + t1r;
        xi[0] = xi[0] + t1i;
        xr[1] = t2r - m * di;
        xi[1] = t2i + m * dr;
        xr[2] = t2r + m * di;
        xi[2] = t2i - m * dr;
    } else {  // R == 5
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = s * T(0.95105651629515357212), s2 = s * T(0.58778525229247312917);
        T b1r = xr[1] + xr[4], b1i = xi[1] + xi[4], b2r = xr[2] + xr[3], b2i = xi[2] + xi[3];
        T d1r = xr[1] - xr[4], d1i = xi[1] - xi[4], d2r = xr[2] - xr[3], d2i = xi[2] - xi[3];
        T a0r = xr[0], a0i = xi[0];
        T p1r = a0r + c1 * b1r + c2 * b2r, p1i = a0i + c1 * b1i + c2 * b2i;
        T p2r = a0r + c2 * b1r + c1 * b2r, p2i = a0i + c2 * b1i + c1 * b2i;
        T q1r = s1 * d1r + s2 * d2r, q1i = s1 * d1i + s2 * d2i;  // multiplied by i below
        T q2r = s2 * d1r - s1 * d2r, q2i = s2 * d1i - s1 * d2i;
        xr[0] = a0r + b1r + b2r;
        xi[0] = a0i + b1i + b2i;
        xr[1] = p1r - q1i;
        xi[1] = p1i + q1r;
        xr[4] = p1r + q1i;
        xi[4] = p1i - q1r;
        xr[2] = p2r - q2i;
        xi[2] = p2i + q2r;
        xr[3] = p2r + q2i;
        xi[3] = p2i - q2r;
    }
}

// in-register DFT of length 8: two length-4 DFTs of the even and odd inputs, combined with
// w^k = exp(sign i pi k / 4)
template <typename T>
__device__ __forceinline__ void butterfly8(T *xr, T *xi, int sign) {
    T er[4] = {xr[0], xr[2], xr[4], xr[6]}, ei[4] = {xi[0], xi[2], xi[4], xi[6]};
    T orr[4] = {xr[1], xr[3], xr[5], xr[7]}, oi[4] = {xi[1], xi[3], xi[5], xi[7]};
    butterfly(4, er, ei, sign);
    butterfly(4, orr, oi, sign);
    const T h = T(0.70710678118654752440), s = T(sign);
    // w^1 = h (1 + s i), w^2 = s i, w^3 = h (-1 + s i)
    T t1r = h * (orr[1] - s * oi[1]), t1i = h * (oi[1] + s * orr[1]);
    T t2r = -s * oi[2], t2i = s * orr[2];
    T t3r = h * (-orr[3] - s * oi[3]), t3i = h * (-oi[3] + s * orr[3]);
    xr[0] = er[0] + orr[0];
    xi[0] = ei[0] + oi[0];
    xr[4] = er[0] - orr[0];
    xi[4] = ei[0] - oi[0];
    xr[1] = er[1] + t1r;
    xi[1] = ei[1] + t1i;
    xr[5] = er[1] - t1r;
    xi[5] = ei[1] - t1i;
    xr[2] = er[2] + t2r;
    xi[2] = ei[2] + t2i;
    xr[6] = er[2] - t2r;
    xi[6] = ei[2] - t2i;
    xr[3] = er[3] + t3r;
    xi[3] = ei[3] + t3i;
    xr[7] = er[3] - t3r;
    xi[7] = ei[3] - t3i;
}

// Output q of butterfly jj of a Stockham pass of prime radix R > 5, summed directly: the butterfly's
// inputs x_r (element jj + r n/R), each twiddled by W^(r jm stride), times omega_R^(r q) = W^(r q n/R),
// i.e. sum_r x_r W^(r E mod n) with E = jm stride + q n/R.  The R-term sum runs in FP64 whatever the
// storage type (its rounding grows with R, where a factored transform's grows with log R).  `at(e)` is
// the twiddle W^e as (cos, -sin) of 2 pi e / n; `in(m)` returns element m of the line.
template <typename In, typename Tw>
__device__ __forceinline__ void dft_radix(int n, int R, int jj, int q, int Ns, int stride, int sign, In in, Tw at, double &yr,
                                          double &yi) {
    const int nbf = n / R, jm = jj % Ns;
    const int E = (int)(((long long)jm * stride + (long long)q * nbf) % n);
    double ar = 0.0, ai = 0.0;
    int e = 0;
    for (int r = 0; r < R; r++) {
        double xr, xi, wr, wi;
        in(jj + r * nbf, xr, xi);
        at(e, wr, wi);
        if (sign > 0) wi = -wi;
        ar += xr * wr - xi * wi;
        ai += xr * wi + xi * wr;
        e += E;
        if (e >= n) e -= n;
    }
    yr = ar;
    yi = ai;
}

// Stockham DFT of kc interleaved batches held in LDS as [n][kc] (re and im planes); ping-pongs
// between buffer 0 and 1 and returns the index of the buffer holding the result.  Radices 2 to 5 take
// the butterflies; a prime radix above 5 takes the direct sums of dft_radix, one output per work item.
template <typename T>
__device__ int lds_fft(T *re[2], T *im[2], const Radix &rad, int kc, const double *tw, int sign) {
    const int n = rad.n;
    int cur = 0, Ns = 1;
    for (int s = 0; s < rad.nr; s++) {
        const int R = rad.r[s];
        const int nb = n / R;
        const int stride = n / (Ns * R);
        T *ir = re[cur], *ii = im[cur], *orr = re[cur ^ 1], *oi = im[cur ^ 1];
        if (R > 5) {
            for (int idx = threadIdx.x; idx < n * kc; idx += blockDim.x) {
                const int b = idx % kc, o = idx / kc, jj = o % nb, q = o / nb;
                double yr, yi;
                dft_radix(
                    n, R, jj, q, Ns, stride, sign, [&](int m, double &xr, double &xi) { xr = double(ir[m * kc + b]); xi = double(ii[m * kc + b]); },
                    [&](int e, double &wr, double &wi) { wr = tw[2 * e]; wi = tw[2 * e + 1]; }, yr, yi);
                const int od = (jj / Ns) * Ns * R + jj % Ns;
                orr[(od + q * Ns) * kc + b] = T(yr);
                oi[(od + q * Ns) * kc + b] = T(yi);
            }
            __syncthreads();
            cur ^= 1;
            Ns *= R;
            continue;
        }
        for (int idx = threadIdx.x; idx < nb * kc; idx += blockDim.x) {
            int b = idx % kc, jj = idx / kc;
            T xr[5], xi[5];
            int jm = jj % Ns;
#pragma unroll
            for (int r = 0; r < 5; r++) {
                if (r < R) {
                    xr[r] = ir[(jj + r * nb) * kc + b];
                    xi[r] = ii[(jj + r * nb) * kc + b];
                }
            }
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < 5; r++) {
                    if (r < R) {
                        T wr, wi;
                        twiddle(tw, r * jm * stride, sign, wr, wi);
                        cmul(xr[r], xi[r], wr, wi);
                    }
                }
            }
            butterfly(R, xr, xi, sign);
            int od = (jj / Ns) * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < 5; r++) {
                if (r < R) {
                    orr[(od + r * Ns) * kc + b] = xr[r];
                    oi[(od + r * Ns) * kc + b] = xi[r];
                }
            }
        }
        __syncthreads();
        cur ^= 1;
        Ns *= R;
    }
    return cur;
}

#define IX(i, j, k) ((long long)(i) * g.NYNZ + (long long)(j) * g.NZ + (long long)(k))

// P1: divergence3d_periodic prologue + forward DFT along i; block = (j, level chunk)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_div_fft_x(Geo g, Cols c, const T *__restrict__ UP, const T *__restrict__ VP,
                                                        const T *__restrict__ WP, T *__restrict__ spec, Radix rad, const double *tw,
                                                        int nk2, int kc, double step) {
    extern __shared__ unsigned char smem[];
    const int n = rad.n;
    T *re[2] = {(T *)smem, (T *)smem + 2 * n * kc};
    T *im[2] = {(T *)smem + n * kc, (T *)smem + 3 * n * kc};
    const int j = blockIdx.x;
    const int kk0 = blockIdx.y * kc;
    const int jj = j == 0 ? 1 : (j == g.NY - 1 ? g.NY - 2 : j);  // mirrored rows, pressure.cpp:1408-1415
    const double sdt = step * g.dt;
    // element idx = i*kc + b; when kc divides the block size each thread keeps its level b
    const bool fixed_b = (blockDim.x % kc) == 0;
    const int di = blockDim.x / kc;
    for (int idx = threadIdx.x, b = threadIdx.x % kc, i = threadIdx.x / kc - di; idx < n * kc; idx += blockDim.x) {
        i += di;
        if (!fixed_b) {
            b = idx % kc;
            i = idx / kc;
        }
        int kk = kk0 + b;
        T d = T(0);
        if (kk < nk2) {
            // the periodic column UP(NX-3) is read as UP(3), the value periodic_uvw_boundaries puts there
            int k = kk + 1, ia = i + 3;
            int ie = ia + 1 == g.NX - 3 ? 3 : ia + 1;
            T rhou = T(c.rhou[k]);
            d = rhou * (UP[IX(ie, jj, k)] - UP[IX(ia, jj, k)]) * T(g.one_d_dx) +
                rhou * (VP[IX(ia, jj + 1, k)] - VP[IX(ia, jj, k)]) * T(g.one_d_dy) +
                (T(c.rhow[k + 1]) * WP[IX(ia, jj, k + 1)] - T(c.rhow[k]) * WP[IX(ia, jj, k)]) * (T(g.one_d_dz) * T(c.mu[k]));
            d /= T(sdt);
        }
        re[0][idx] = d;
        im[0][idx] = T(0);
    }
    __syncthreads();
    int out = lds_fft(re, im, rad, kc, tw, -1);
    // spec[kx][j][kk] (interleaved complex)
    for (int idx = threadIdx.x; idx < n * kc; idx += blockDim.x) {
        int b = idx % kc, kx = idx / kc, kk = kk0 + b;
        if (kk < nk2) {
            long long o = (((long long)kx * g.NY + j) * nk2 + kk) * 2;
            spec[o] = re[out][idx];
            spec[o + 1] = im[out][idx];
        }
    }
}

// P2/P4: DFT along j for fixed kx; block = (kx, level chunk)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_fft_y(T *__restrict__ spec, int nj, int nk2, Radix rad, const double *tw, int kc, int sign) {
    extern __shared__ unsigned char smem[];
    const int n = rad.n;
    T *re[2] = {(T *)smem, (T *)smem + 2 * n * kc};
    T *im[2] = {(T *)smem + n * kc, (T *)smem + 3 * n * kc};
    const int kx = blockIdx.x, kk0 = blockIdx.y * kc;
    for (int idx = threadIdx.x; idx < n * kc; idx += blockDim.x) {
        int b = idx % kc, ky = idx / kc, kk = kk0 + b;
        T vr = T(0), vi = T(0);
        if (kk < nk2) {
            long long o = (((long long)kx * nj + ky) * nk2 + kk) * 2;
            vr = spec[o];
            vi = spec[o + 1];
        }
        re[0][idx] = vr;
        im[0][idx] = vi;
    }
    __syncthreads();
    int out = lds_fft(re, im, rad, kc, tw, sign);
    for (int idx = threadIdx.x; idx < n * kc; idx += blockDim.x) {
        int b = idx % kc, ky = idx / kc, kk = kk0 + b;
        if (kk < nk2) {
            long long o = (((long long)kx * nj + ky) * nk2 + kk) * 2;
            spec[o] = re[out][idx];
            spec[o + 1] = im[out][idx];
        }
    }
}

// Tridiagonal multipliers F2 of pressure.cpp:840-855 for every (kx,ky): a pure function of the
// coefficients, so they are formed once per context instead of once per solve.
__global__ void k_thomas_factors(int ni, int nj, int nk, const double *__restrict__ coskx, const double *__restrict__ cosly,
                                 const double *__restrict__ rhou, const double *__restrict__ rhow, const double *__restrict__ mu,
                                 const double *__restrict__ mw, const double *__restrict__ acoef, const double *__restrict__ ccoef,
                                 double dx, double dz, double *__restrict__ ffac, double *__restrict__ zpiv) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ni * nj) return;
    int kx = q / nj, ky = q % nj;
    int nk2 = nk - 2;
    // layout [kx][kk][ky]: the columns of one kx are adjacent for every level
    double *F = ffac + (long long)kx * nk2 * nj + ky;
    double *Z = zpiv + (long long)kx * nk2 * nj + ky;
    const long long fs = nj;
    double ck = coskx[kx], cl = cosly[ky];
    // B2 of pressure.cpp:787-791
    double B = rhou[1] * 2.0 * (ck + cl - 2.0) / (dx * dx) - mu[1] * mw[2] * rhow[2] / (dz * dz);
    double z = 1.0 / B;
    double Fp = ccoef[1] * z;
    F[0] = Fp;
    Z[0] = z;
    for (int k = 2; k < nk - 1; k++) {
        if (k < nk - 2) B = rhou[k] * 2.0 * (ck + cl - 2.0) / (dx * dx) - mu[k] * (mw[k + 1] * rhow[k + 1] + mw[k] * rhow[k]) / (dz * dz);
        else B = rhou[nk - 2] * 2.0 * (ck + cl - 2.0) / (dx * dx) - (mu[nk - 2] * mw[nk - 2] * rhow[nk - 2]) / (dz * dz);
        z = 1.0 / (B - acoef[k] * Fp);
        Fp = ccoef[k] * z;
        F[(k - 1) * fs] = Fp;
        Z[(k - 1) * fs] = z;
    }
}

// P3: normalisation (pressure.cpp:824-827), forward elimination (837-856), mean-mode gauge (858-859)
// and back substitution (864-871), in FP64 whatever the storage precision.
template <typename T>
__global__ void k_thomas(T *__restrict__ spec, int ni, int nj, int nk, const double *__restrict__ coskx, const double *__restrict__ cosly,
                         const double *__restrict__ rhou, const double *__restrict__ rhow, const double *__restrict__ mu,
                         const double *__restrict__ mw, const double *__restrict__ acoef, const double *__restrict__ ffac, double dx,
                         double dz) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ni * nj) return;
    int kx = q / nj, ky = q % nj;
    const int nk2 = nk - 2;
    T *x = spec + (long long)q * nk2 * 2;
    const long long fs = nj;
    const double *F = ffac + (long long)kx * nk2 * nj + ky;
    const double size2d = (double)(ni * nj);
    double ck = coskx[kx], cl = cosly[ky];
    double B = rhou[1] * 2.0 * (ck + cl - 2.0) / (dx * dx) - mu[1] * mw[2] * rhow[2] / (dz * dz);
    double z = 1.0 / B;
    double pr = (double(x[0]) / size2d) * z, pi = (double(x[1]) / size2d) * z;
    x[0] = T(pr);
    x[1] = T(pi);
    double Fp = F[0];
    for (int k = 2; k < nk - 1; k++) {
        if (k < nk - 2) B = rhou[k] * 2.0 * (ck + cl - 2.0) / (dx * dx) - mu[k] * (mw[k + 1] * rhow[k + 1] + mw[k] * rhow[k]) / (dz * dz);
        else B = rhou[nk - 2] * 2.0 * (ck + cl - 2.0) / (dx * dx) - (mu[nk - 2] * mw[nk - 2] * rhow[nk - 2]) / (dz * dz);
        z = 1.0 / (B - acoef[k] * Fp);
        double xr = double(x[2 * (k - 1)]) / size2d, xi = double(x[2 * (k - 1) + 1]) / size2d;
        pr = (xr - acoef[k] * pr) * z;
        pi = (xi - acoef[k] * pi) * z;
        x[2 * (k - 1)] = T(pr);
        x[2 * (k - 1) + 1] = T(pi);
        Fp = F[(k - 1) * fs];
    }
    if (kx == 0 && ky == 0) {  // OUT3D(0,0,NZ-2) = 0
        pr = 0.0;
        pi = 0.0;
        x[2 * (nk2 - 1)] = T(0);
        x[2 * (nk2 - 1) + 1] = T(0);
    }
    for (int k = nk - 3; k >= 1; k--) {
        double f = F[(k - 1) * fs];
        pr = double(x[2 * (k - 1)]) - f * pr;
        pi = double(x[2 * (k - 1) + 1]) - f * pi;
        x[2 * (k - 1)] = T(pr);
        x[2 * (k - 1) + 1] = T(pi);
    }
}

// P5: backward DFT along i, real part -> PI (pressure.cpp:878-892) + periodic halos (boundaries.cpp:533-540)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_ifft_x_out(Geo g, const T *__restrict__ spec, T *__restrict__ PI, Radix rad,
                                                         const double *tw, int nk2, int kc) {
    extern __shared__ unsigned char smem[];
    const int n = rad.n;
    T *re[2] = {(T *)smem, (T *)smem + 2 * n * kc};
    T *im[2] = {(T *)smem + n * kc, (T *)smem + 3 * n * kc};
    const int j = blockIdx.x, kk0 = blockIdx.y * kc;
    for (int idx = threadIdx.x; idx < n * kc; idx += blockDim.x) {
        int b = idx % kc, kx = idx / kc, kk = kk0 + b;
        T vr = T(0), vi = T(0);
        if (kk < nk2) {
            long long o = (((long long)kx * g.NY + j) * nk2 + kk) * 2;
            vr = spec[o];
            vi = spec[o + 1];
        }
        re[0][idx] = vr;
        im[0][idx] = vi;
    }
    __syncthreads();
    int out = lds_fft(re, im, rad, kc, tw, +1);
    const int shift = g.NX - 6;
    const bool fixed_b = (blockDim.x % kc) == 0;
    const int di = blockDim.x / kc;
    for (int idx = threadIdx.x, b = threadIdx.x % kc, i = threadIdx.x / kc - di; idx < n * kc; idx += blockDim.x) {
        i += di;
        if (!fixed_b) {
            b = idx % kc;
            i = idx / kc;
        }
        const int kk = kk0 + b;
        if (kk >= nk2) continue;
        T v = re[out][idx];
        int ia = i + 3;
        int k = kk + 1;
        int ks[3] = {k, -1, -1};
        if (kk == 0) ks[1] = 0;
        if (kk == nk2 - 1) ks[2] = g.NZ - 1;
        for (int t = 0; t < 3; t++) {
            if (ks[t] < 0) continue;
            PI[IX(ia, j, ks[t])] = v;
            if (ia >= shift && ia < shift + 3) PI[IX(ia - shift, j, ks[t])] = v;
            if (ia >= 3 && ia < 6) PI[IX(ia + shift, j, ks[t])] = v;
        }
    }
}


// ---- half-spectrum path (even ni) ------------------------------------------------------------------
// The divergence is real, so two levels travel as the real and imaginary parts of one complex DFT along
// i and only kx = 0..ni/2 is kept (the rest is the conjugate mirror).  After the j-transform and the
// tridiagonal solve the real part of the inverse is rebuilt from that half, which is what taking the real
// part of the full inverse (pressure.cpp:881-890) computes.

// in-place Stockham DFT over LDS: element (n, b) at re[n*sn + b*sb], im[...]; every butterfly's inputs are
// read into registers before a barrier, then written to their output slots
// In-place Stockham DFT of nbatch sequences of length n held in LDS (element (jj, b) at jj*sn + b*sb, re
// and im in separate planes); twiddles come from an LDS copy of the table (twl[2e], twl[2e+1] = cos,
// -sin of 2 pi e / n).  Work item idx = tid + t*nthreads is butterfly jj = idx / nbatch of batch
// b = idx % nbatch; that split is done once, and the per-pass index arithmetic uses shifts while the
// running length Ns is a power of two (every pass of the radix-4/2-first plans before a 3 or 5).
// One radix-R pass of the power-of-two plans of lds_fft_ip with R a template parameter: items past the
// end read a valid slot (and are not written back), so the pass is straight-line code and every LDS
// load is issued before the first is waited on.
template <typename T, int R, int MAXB>
__device__ __forceinline__ void fft_pass_pow2(T *re, T *im, int n, int Ns, int lg, int sn, int sb, const int *bq, const int *jq,
                                              const T *twl, int sign) {
    const int nbf = n / R;
    const int stride = n / (Ns * R);
    T xr[MAXB][R], xi[MAXB][R];
    int oidx[MAXB];
#pragma unroll
    for (int t = 0; t < MAXB; t++) {
        const bool ok = jq[t] < nbf;
        const int jj = ok ? jq[t] : 0, b = bq[t];
        const int jm = jj & (Ns - 1), jd = jj >> lg;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int e = (jj + r * nbf) * sn + b * sb;
            xr[t][r] = re[e];
            xi[t][r] = im[e];
        }
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const int e = r * jm * stride;
                const T wr = twl[2 * e], wi = sign < 0 ? twl[2 * e + 1] : -twl[2 * e + 1];
                cmul(xr[t][r], xi[t][r], wr, wi);
            }
        }
        if constexpr (R == 8) butterfly8(xr[t], xi[t], sign);
        else butterfly(R, xr[t], xi[t], sign);
        oidx[t] = ok ? (jd * Ns * R + jm) * sn + b * sb : -1;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXB; t++) {
        if (oidx[t] >= 0) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                re[oidx[t] + r * Ns * sn] = xr[t][r];
                im[oidx[t] + r * Ns * sn] = xi[t][r];
            }
        }
    }
    __syncthreads();
}

// butterflies a thread holds in a radix-8 pass of a plan whose other passes hold MAXB (the radix-8 pass
// has half the radix-4 count; a third of MAXB keeps its registers within what the radix-4 passes use)
constexpr int pow2_maxb8(int maxb) { return maxb / 3 > 0 ? maxb / 3 : 1; }

// the power-of-two plans of lds_fft_ip alone (radix-4 passes, with one radix-2 or radix-8 pass when
// log2 n is odd)
template <typename T, int MAXB>
__device__ void lds_fft_pow2(T *re, T *im, const Radix &rad, int nbatch, int sn, int sb, const T *twl, int sign) {
    const int n = rad.n;
    int bq[MAXB], jq[MAXB];
#pragma unroll
    for (int t = 0; t < MAXB; t++) {
        const int idx = threadIdx.x + t * (int)blockDim.x;
        bq[t] = idx % nbatch;
        jq[t] = idx / nbatch;
    }
    int Ns = 1, lg = 0;
    for (int s = 0; s < rad.nr; s++) {
        if (rad.r[s] == 4) fft_pass_pow2<T, 4, MAXB>(re, im, n, Ns, lg, sn, sb, bq, jq, twl, sign);
        else if (rad.r[s] == 8) fft_pass_pow2<T, 8, pow2_maxb8(MAXB)>(re, im, n, Ns, lg, sn, sb, bq, jq, twl, sign);
        else fft_pass_pow2<T, 2, MAXB>(re, im, n, Ns, lg, sn, sb, bq, jq, twl, sign);
        Ns *= rad.r[s];
        lg = 31 - __clz(Ns);
    }
}

// power-of-two plan of the half-slab j pass: radix-4 passes, the first one radix 8 when log2 n is odd
// (so no radix-2 pass, whose butterflies would be twice as many per thread)
bool factorize(int n, Radix &rad);

inline bool pow2_plan48(int n, Radix &rad, bool radix8) {
    if (n < 2 || (n & (n - 1)) != 0) return false;
    if (!radix8) return factorize(n, rad);  // 4, ..., 4, 2
    rad.n = n;
    rad.nr = 0;
    int m = n;
    if (__builtin_ctz(n) % 2 == 1) {
        if (n == 2) {
            rad.r[rad.nr++] = 2;
            return true;
        }
        rad.r[rad.nr++] = 8;
        m /= 8;
    }
    while (m > 1 && rad.nr < kMaxRad) {
        rad.r[rad.nr++] = 4;
        m /= 4;
    }
    return m == 1;
}

template <typename T, int MAXB, bool POW2 = false>
__device__ __forceinline__ void lds_fft_ip(T *re, T *im, const Radix &rad, int nbatch, int sn, int sb, const T *twl, int sign) {
    const int n = rad.n;
    const int nthr = blockDim.x;
    int bq[MAXB], jq[MAXB];
#pragma unroll
    for (int t = 0; t < MAXB; t++) {
        const int idx = threadIdx.x + t * nthr;
        bq[t] = idx % nbatch;
        jq[t] = idx / nbatch;
    }
    if (POW2 && (n & (n - 1)) == 0) {  // power of two: radix-4 passes, then one radix-2 pass when log2 n is odd
        int Ns = 1, lg = 0;
        for (int s = 0; s < rad.nr; s++) {
            if (rad.r[s] == 4) fft_pass_pow2<T, 4, MAXB>(re, im, n, Ns, lg, sn, sb, bq, jq, twl, sign);
            else fft_pass_pow2<T, 2, MAXB>(re, im, n, Ns, lg, sn, sb, bq, jq, twl, sign);
            Ns *= rad.r[s];
            lg = 31 - __clz(Ns);
        }
        return;
    }
    int Ns = 1, lg = 0;
    for (int s = 0; s < rad.nr; s++) {
        const int R = rad.r[s];
        const int nbf = n / R;
        const int stride = n / (Ns * R);
        const bool pow2 = (Ns & (Ns - 1)) == 0;
        T xr[MAXB][5], xi[MAXB][5];
        int oidx[MAXB];
#pragma unroll
        for (int t = 0; t < MAXB; t++) {
            const int jj = jq[t], b = bq[t];
            if (jj < nbf) {
                const int jm = pow2 ? (jj & (Ns - 1)) : jj % Ns;
                const int jd = pow2 ? (jj >> lg) : jj / Ns;
#pragma unroll
                for (int r = 0; r < 5; r++) {
                    if (r < R) {
                        const int e = (jj + r * nbf) * sn + b * sb;
                        xr[t][r] = re[e];
                        xi[t][r] = im[e];
                    }
                }
                if (Ns > 1) {
#pragma unroll
                    for (int r = 1; r < 5; r++) {
                        if (r < R) {
                            const int e = r * jm * stride;
                            const T wr = twl[2 * e], wi = sign < 0 ? twl[2 * e + 1] : -twl[2 * e + 1];
                            cmul(xr[t][r], xi[t][r], wr, wi);
                        }
                    }
                }
                butterfly(R, xr[t], xi[t], sign);
                oidx[t] = (jd * Ns * R + jm) * sn + b * sb;
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < MAXB; t++) {
            if (jq[t] < nbf) {
#pragma unroll
                for (int r = 0; r < 5; r++) {
                    if (r < R) {
                        re[oidx[t] + r * Ns * sn] = xr[t][r];
                        im[oidx[t] + r * Ns * sn] = xi[t][r];
                    }
                }
            }
        }
        __syncthreads();
        Ns *= R;
        if ((Ns & (Ns - 1)) == 0) lg = 31 - __clz(Ns);
    }
}

// copy the (cos, -sin) table of length n into LDS in the storage type
template <typename T>
__device__ __forceinline__ void load_twiddles(T *twl, const double *tw, int n) {
    for (int e = threadIdx.x; e < 2 * n; e += blockDim.x) twl[e] = T(tw[e]);
}

// j transforms of the three-kernel j pass for power-of-two nj: kc levels of one kx in LDS as [ky][level]
// (one padding level), transformed in place by the straight-line radix-8/4 passes with the twiddles in
// LDS; the same layout in HBM as k_fft_y
// 512 threads with two butterflies each keep the registers low (two blocks per CU) and still take 8
// FP32 or 8 FP64 levels per block, whole 64- or 128-byte runs of the spectrum per ky
constexpr int kFftYThreads = 512, kFftYMaxB = 2;
constexpr size_t kFftYLds = 96 * 1024;
template <typename T, int MAXB>
__global__ void __launch_bounds__(kFftYThreads) k_fft_y_pow2(T *__restrict__ spec, int nj, int nk2, Radix rad, const double *tw, int kc,
                                                         int sign) {
    extern __shared__ unsigned char smem[];
    const int n = rad.n, P2 = kc + 1;
    T *re = (T *)smem, *im = re + n * P2, *twl = im + n * P2;
    load_twiddles(twl, tw, n);
    const int kx = blockIdx.x, kk0 = blockIdx.y * kc;
    for (int idx = threadIdx.x; idx < n * kc; idx += blockDim.x) {
        const int b = idx % kc, ky = idx / kc, kk = kk0 + b;
        T vr = T(0), vi = T(0);
        if (kk < nk2) {
            const long long o = (((long long)kx * nj + ky) * nk2 + kk) * 2;
            vr = spec[o];
            vi = spec[o + 1];
        }
        re[ky * P2 + b] = vr;
        im[ky * P2 + b] = vi;
    }
    __syncthreads();
    lds_fft_pow2<T, MAXB>(re, im, rad, kc, P2, 1, twl, sign);
    for (int idx = threadIdx.x; idx < n * kc; idx += blockDim.x) {
        const int b = idx % kc, ky = idx / kc, kk = kk0 + b;
        if (kk < nk2) {
            const long long o = (((long long)kx * nj + ky) * nk2 + kk) * 2;
            spec[o] = re[ky * P2 + b];
            spec[o + 1] = im[ky * P2 + b];
        }
    }
}
inline size_t fft_y_pow2_lds(int n, int kc, size_t elt) { return 2 * (size_t)n * (kc + 1) * elt + 2 * (size_t)n * elt; }
// levels per block of k_fft_y_pow2 for the plan `rad` (the most that fit the LDS budget and the
// butterflies a thread holds in every pass), 0 when none do
template <typename T>
int fft_y_pow2_chunk(const Radix &rad, int nk2) {
    const int n = rad.n, maxb = kFftYMaxB;
    for (int kc = nk2 < 32 ? nk2 : 32; kc >= 1; kc--) {
        if (fft_y_pow2_lds(n, kc, sizeof(T)) > kFftYLds) continue;
        bool ok = true;
        for (int st = 0; ok && st < rad.nr; st++) {
            const int items = n / rad.r[st] * kc, per = (items + kFftYThreads - 1) / kFftYThreads;
            ok = per <= (rad.r[st] == 8 ? pow2_maxb8(maxb) : maxb);
        }
        if (ok) return kc;
    }
    return 0;
}

constexpr int kMaxBX = 4;   // butterflies per thread in the x passes (register-resident)
constexpr int kMaxBY = 4;   // and in the fused y pass (FP32, 1024 threads)
constexpr int kMaxBY64 = 6; // the fused y pass in FP64 runs 512 threads with up to 6 butterflies each
template <typename T>
constexpr int ypass_threads() { return sizeof(T) == 4 ? 1024 : 512; }
template <typename T>
constexpr int ypass_maxb() { return sizeof(T) == 4 ? kMaxBY : kMaxBY64; }
// the half-slab j pass (power-of-two plans, radix 4 and one radix-8 pass): FP32 1024 threads with up to
// 4 butterflies each, FP64 512 threads with up to 6 (what fits each thread's registers)
template <typename T>
constexpr int half_threads() { return sizeof(T) == 4 ? 1024 : 512; }
template <typename T>
constexpr int half_maxb() { return sizeof(T) == 4 ? 4 : 6; }

// Q1: divergence3d_periodic prologue + forward real DFT along i; block = (j, chunk of kc levels, kc even).
// The periodic column UP(NX-3) is read as UP(3), the value periodic_uvw_boundaries puts there.
// With src != nullptr the transform reads a ready real field src(i+3, j, k) instead of the divergence
// (the transposed solve of the adjoint).
template <typename T>
__device__ __forceinline__ void div_rfft_x_block(const Geo &g, const Cols &c, const T *__restrict__ UP, const T *__restrict__ VP,
                                                 const T *__restrict__ WP, T *__restrict__ spec, const Radix &rad, const double *tw,
                                                 int nk2, int kc, double step, const T *__restrict__ src, int j, int chunk,
                                                 unsigned char *smem) {
    const int n = rad.n, P = kc / 2, nkx = n / 2 + 1;
    T *re = (T *)smem, *im = re + n * P, *twl = im + n * P;
    const int kk0 = chunk * kc;
    load_twiddles(twl, tw, n);
    const int jj = j == 0 ? 1 : (j == g.NY - 1 ? g.NY - 2 : j);  // mirrored rows, pressure.cpp:1408-1415
    const double sdt = step * g.dt;
    // the column coefficients of the block's kc levels (rhou(k), rhow(k+1), rhow(k), mu(k)), read once
    // into LDS instead of once per element
    double *colc = (double *)(twl + 2 * n);
    if (!src) {
        for (int b = (int)threadIdx.x; b < kc; b += (int)blockDim.x) {
            const int k = min(kk0 + b, nk2 - 1) + 1;
            colc[4 * b] = c.rhou[k];
            colc[4 * b + 1] = c.rhow[k + 1];
            colc[4 * b + 2] = c.rhow[k];
            colc[4 * b + 3] = c.mu[k];
        }
        __syncthreads();
    }
    // element idx = i*kc + b.  kU elements per thread per round: every global load of a round is issued
    // (clamped levels, so no load sits behind a branch) before any of them is used; the source-field form
    // and the divergence form are separate loops for the same reason.
    constexpr int kU = 4;
    const int ne = n * kc, nt = (int)blockDim.x;
    auto gather = [&](auto from_src) {
        constexpr bool kSrc = decltype(from_src)::value;
        for (int idx0 = (int)threadIdx.x; idx0 < ne; idx0 += kU * nt) {
            T f0[kU], f1[kU], f2[kU], f3[kU], f4[kU], f5[kU];
            double ru[kU], rw1[kU], rw0[kU], mk[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int idx = min(idx0 + u * nt, ne - 1), b = idx % kc, i = idx / kc;
                const int k = min(kk0 + b, nk2 - 1) + 1, ia = i + 3;
                if (kSrc) {
                    f0[u] = src[IX(ia, j, k)];
                } else {
                    const int ie = ia + 1 == g.NX - 3 ? 3 : ia + 1;
                    f0[u] = UP[IX(ie, jj, k)];
                    f1[u] = UP[IX(ia, jj, k)];
                    f2[u] = VP[IX(ia, jj + 1, k)];
                    f3[u] = VP[IX(ia, jj, k)];
                    f4[u] = WP[IX(ia, jj, k + 1)];
                    f5[u] = WP[IX(ia, jj, k)];
                    ru[u] = colc[4 * b];
                    rw1[u] = colc[4 * b + 1];
                    rw0[u] = colc[4 * b + 2];
                    mk[u] = colc[4 * b + 3];
                }
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int idx = idx0 + u * nt, b = idx % kc, i = idx / kc;
                const int kk = kk0 + b;
                T d = T(0);
                if (kk < nk2 && kSrc) {
                    d = f0[u];
                } else if (kk < nk2) {
                    T rhou = T(ru[u]);
                    d = rhou * (f0[u] - f1[u]) * T(g.one_d_dx) + rhou * (f2[u] - f3[u]) * T(g.one_d_dy) +
                        (T(rw1[u]) * f4[u] - T(rw0[u]) * f5[u]) * (T(g.one_d_dz) * T(mk[u]));
                    d /= T(sdt);
                }
                if (idx < ne) {
                    if (b & 1) im[i * P + (b >> 1)] = d;
                    else re[i * P + (b >> 1)] = d;
                }
            }
        }
    };
    if (src) gather(std::true_type{});
    else gather(std::false_type{});
    __syncthreads();
    lds_fft_ip<T, kMaxBX>(re, im, rad, P, P, 1, twl, -1);
    // split the packed pair: A = (Z[m] + conj Z[-m]) / 2, B = (Z[m] - conj Z[-m]) / 2i
    for (int idx = threadIdx.x; idx < nkx * P; idx += blockDim.x) {
        int pp = idx % P, m = idx / P;
        int mm = m == 0 ? 0 : n - m;
        T zr = re[m * P + pp], zi = im[m * P + pp], wr = re[mm * P + pp], wi = im[mm * P + pp];
        int ka = kk0 + 2 * pp, kb = ka + 1;
        long long o = ((long long)m * g.NY + j) * nk2;
        if (kb < nk2 && (nk2 & 1) == 0) {  // the pair's four values in one store
            st4(spec + (o + ka) * 2, T(0.5) * (zr + wr), T(0.5) * (zi - wi), T(0.5) * (zi + wi), T(0.5) * (wr - zr));
        } else {
            if (ka < nk2) {
                spec[(o + ka) * 2] = T(0.5) * (zr + wr);
                spec[(o + ka) * 2 + 1] = T(0.5) * (zi - wi);
            }
            if (kb < nk2) {
                spec[(o + kb) * 2] = T(0.5) * (zi + wi);
                spec[(o + kb) * 2 + 1] = T(0.5) * (wr - zr);
            }
        }
    }
}

// Q3: backward real DFT along i -> PI (pressure.cpp:878-892) with the ghost levels and periodic halos
template <typename T>
__device__ __forceinline__ void irfft_x_out_block(const Geo &g, const T *__restrict__ spec, T *__restrict__ PI, const Radix &rad,
                                                  const double *tw, int nk2, int kc, int j, int chunk, unsigned char *smem,
                                                  int jhalf = 0) {
    const int n = rad.n, P = kc / 2, half = n / 2;
    T *re = (T *)smem, *im = re + n * P, *twl = im + n * P;
    const int kk0 = chunk * kc;
    load_twiddles(twl, tw, n);
    // rebuild Z = A + iB for m in [0, n) from the half spectra of the two levels of each pair.  kU
    // elements per thread per round: every global load of a round is issued (clamped levels, so no load
    // sits behind a branch) before any of them is used.
    // With jhalf the half-slab j pass left row j of the full inverse as v0(j') + s v1(j'), j' = j mod
    // jhalf, with v0 in row j' and v1 in row jhalf + j' (s = +1 on the first half of the rows, -1 on the
    // second).
    constexpr int kU = 4;
    const int ne = n * P, nt = (int)blockDim.x;
    const int jp = jhalf ? (j < jhalf ? j : j - jhalf) : j;
    const T sg = j < jhalf ? T(1) : T(-1);
    for (int idx0 = (int)threadIdx.x; idx0 < ne; idx0 += kU * nt) {
        T x[kU][4], y[kU][4];  // (re, im) at ka, then at kb: row jp (x) and, with jhalf, row jhalf + jp (y)
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int idx = min(idx0 + u * nt, ne - 1), pp = idx % P, m = idx / P;
            const int ka = min(kk0 + 2 * pp, nk2 - 1), kb = min(kk0 + 2 * pp + 1, nk2 - 1);
            const int ms = m <= half ? m : n - m;
            const long long o0 = ((long long)ms * g.NY + jp) * nk2;
            x[u][0] = spec[(o0 + ka) * 2];
            x[u][1] = spec[(o0 + ka) * 2 + 1];
            x[u][2] = spec[(o0 + kb) * 2];
            x[u][3] = spec[(o0 + kb) * 2 + 1];
            if (jhalf) {
                const long long o1 = ((long long)ms * g.NY + jhalf + jp) * nk2;
                y[u][0] = spec[(o1 + ka) * 2];
                y[u][1] = spec[(o1 + ka) * 2 + 1];
                y[u][2] = spec[(o1 + kb) * 2];
                y[u][3] = spec[(o1 + kb) * 2 + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int idx = idx0 + u * nt, pp = idx % P, m = idx / P;
            const int ka = kk0 + 2 * pp, kb = ka + 1;
            const int ms = m <= half ? m : n - m;
            T ar = T(0), ai = T(0), br = T(0), bi = T(0);
            if (jhalf) {
                if (ka < nk2) {
                    ar = x[u][0] + sg * y[u][0];
                    ai = x[u][1] + sg * y[u][1];
                }
                if (kb < nk2) {
                    br = x[u][2] + sg * y[u][2];
                    bi = x[u][3] + sg * y[u][3];
                }
            } else {
                if (ka < nk2) {
                    ar = x[u][0];
                    ai = x[u][1];
                }
                if (kb < nk2) {
                    br = x[u][2];
                    bi = x[u][3];
                }
            }
            if (ms == 0 || ms == half) {  // self-conjugate wavenumbers: only the real part is kept
                ai = T(0);
                bi = T(0);
            }
            if (m > half) {  // conjugate mirror
                ai = -ai;
                bi = -bi;
            }
            if (idx < ne) {
                re[m * P + pp] = ar - bi;
                im[m * P + pp] = ai + br;
            }
        }
    }
    __syncthreads();
    lds_fft_ip<T, kMaxBX>(re, im, rad, P, P, 1, twl, +1);
    const int shift = g.NX - 6;
    const bool fixed_b = (blockDim.x % kc) == 0;
    const int di = blockDim.x / kc;
    for (int idx = threadIdx.x, b = threadIdx.x % kc, i = threadIdx.x / kc - di; idx < n * kc; idx += blockDim.x) {
        i += di;
        if (!fixed_b) {
            b = idx % kc;
            i = idx / kc;
        }
        const int kk = kk0 + b;
        if (kk >= nk2) continue;
        T v = (b & 1) ? im[i * P + (b >> 1)] : re[i * P + (b >> 1)];
        int ia = i + 3, k = kk + 1;
        int ks[3] = {k, -1, -1};
        if (kk == 0) ks[1] = 0;
        if (kk == nk2 - 1) ks[2] = g.NZ - 1;
        for (int t = 0; t < 3; t++) {
            if (ks[t] < 0) continue;
            PI[IX(ia, j, ks[t])] = v;
            if (ia >= shift && ia < shift + 3) PI[IX(ia - shift, j, ks[t])] = v;
            if (ia >= 3 && ia < 6) PI[IX(ia + shift, j, ks[t])] = v;
        }
    }
}

// B2 of pressure.cpp:787-791 for level k (1..nk-2) of column (ck, cl)
__device__ __forceinline__ double bcoef(int k, int nk, double ck, double cl, const double *rhou, const double *rhow, const double *mu,
                                        const double *mw, double dx, double dz) {
    if (k == 1) return rhou[1] * 2.0 * (ck + cl - 2.0) / (dx * dx) - mu[1] * mw[2] * rhow[2] / (dz * dz);
    if (k < nk - 2) return rhou[k] * 2.0 * (ck + cl - 2.0) / (dx * dx) - mu[k] * (mw[k + 1] * rhow[k + 1] + mw[k] * rhow[k]) / (dz * dz);
    return rhou[nk - 2] * 2.0 * (ck + cl - 2.0) / (dx * dx) - (mu[nk - 2] * mw[nk - 2] * rhow[nk - 2]) / (dz * dz);
}

struct ThomasArgs {
    const double *coskx, *cosly, *rhou, *rhow, *mu, *mw, *acoef, *ffac;
    double dx, dz;
    int ni, nj, nk;
    const double *zpiv;  // pivot reciprocals z_k of every column, same layout as ffac
};

// The sweeps below walk the levels in chunks of kTC and load the next chunk's table values (pivots or
// multipliers, one column of the [kx][kk][ky] tables) while the current chunk is eliminated, so the
// dependent recurrence does not wait on global-memory latency level by level.
constexpr int kThomasChunk = 8;  // levels per chunk of the Thomas sweeps

// Forward elimination and back substitution of one real sequence (the real or the imaginary part) of one
// (kx,ky) column whose values sit at x[kk*xs]; FP64 arithmetic with the pivots and multipliers of the
// tables; the (0,0) column (m0) keeps its eliminated values in the FP64 scratch mean[2*kk] instead of the
// storage type (pressure.cpp:837-871).  Level k = kk+1; acoef[1] = 0, so the first level needs no
// special case.
// The sweeps walk kTC levels at a time, and the table entries of the next chunk are loaded while the
// current chunk is eliminated, into the second of two register sets that take turns: a loop that copied
// the prefetched set into the current one at its end had to wait there for the loads just issued, so
// every chunk waited a whole memory latency.  Every load is unconditional (clamped indices); only the
// last partial chunk guards its steps.  The (0,0) column differs only in where its values are kept (a
// per-lane choice of store and of load): an instance of its own would diverge from the other lanes of
// its wavefront and run the whole chain twice.
template <int kTC>
__device__ __forceinline__ void thomas_tab_up(const double *Tab, long long fs, int k0, int hi, double (&v)[kTC]) {
#pragma unroll
    for (int t = 0; t < kTC; t++) v[t] = Tab[min(k0 + t, hi) * fs];
}
template <int kTC>
__device__ __forceinline__ void thomas_tab_down(const double *Tab, long long fs, int k0, double (&v)[kTC]) {
#pragma unroll
    for (int t = 0; t < kTC; t++) v[t] = Tab[max(k0 - t, 0) * fs];
}
// p_kk = (x_kk / (ni nj) - a_k p_{kk-1}) z_kk for the chunk's levels kk = k0 .. k0+kTC-1 (< N unless FULL)
template <typename T, bool FULL, int kTC>
__device__ __forceinline__ void thomas_fwd_chunk(const ThomasArgs &ta, const double (&z)[kTC], int k0, int N, double size2d, T *x,
                                                 int xs, bool m0, double *mean, double &p) {
    double v[kTC], ac[kTC];
#pragma unroll
    for (int t = 0; t < kTC; t++) {
        const int kk = FULL ? k0 + t : min(k0 + t, N - 1);
        v[t] = double(x[kk * xs]);
        ac[t] = ta.acoef[kk + 1];
    }
#pragma unroll
    for (int t = 0; t < kTC; t++) {
        if (FULL || k0 + t < N) p = (v[t] / size2d - ac[t] * p) * z[t];
        v[t] = p;
    }
#pragma unroll
    for (int t = 0; t < kTC; t++) {
        if (FULL || k0 + t < N) {
            if (m0) mean[2 * (k0 + t)] = v[t];
            else x[(k0 + t) * xs] = T(v[t]);
        }
    }
}
// x_kk = y_kk - F_kk x_{kk+1} for kk = k0 .. k0-kTC+1 (>= 0 unless FULL)
template <typename T, bool FULL, int kTC>
__device__ __forceinline__ void thomas_back_chunk(const double (&f)[kTC], int k0, T *x, int xs, bool m0, const double *mean,
                                                  double &p) {
    double v[kTC];
#pragma unroll
    for (int t = 0; t < kTC; t++) {
        const int kk = FULL ? k0 - t : max(k0 - t, 0);
        const double vm = mean[2 * kk], vx = double(x[kk * xs]);
        v[t] = m0 ? vm : vx;
    }
#pragma unroll
    for (int t = 0; t < kTC; t++) {
        if (FULL || k0 - t >= 0) p = v[t] - f[t] * p;
        v[t] = p;
    }
#pragma unroll
    for (int t = 0; t < kTC; t++)
        if (FULL || k0 - t >= 0) x[(k0 - t) * xs] = T(v[t]);
}
template <typename T, int kTC = kThomasChunk>
__device__ void thomas_part(const ThomasArgs &ta, int kx, int ky, T *x, int xs, double *mean) {
    const int N = ta.nk - 2;
    const long long fs = ta.nj, col = (long long)kx * N * ta.nj + ky;
    const double *F = ta.ffac + col, *Z = ta.zpiv + col;
    const double size2d = (double)(ta.ni * ta.nj);
    const bool m0 = (kx == 0 && ky == 0);
    double A[kTC], B[kTC], p = 0.0;
    // full chunks two at a time (one path through the loop, so the loads of the set not in use stay in
    // flight across its back edge), then the levels left
    thomas_tab_up<kTC>(Z, fs, 0, N - 1, A);
    int k0 = 0;
    for (; k0 + 2 * kTC <= N; k0 += 2 * kTC) {
        thomas_tab_up<kTC>(Z, fs, k0 + kTC, N - 1, B);
        thomas_fwd_chunk<T, true, kTC>(ta, A, k0, N, size2d, x, xs, m0, mean, p);
        thomas_tab_up<kTC>(Z, fs, k0 + 2 * kTC, N - 1, A);
        thomas_fwd_chunk<T, true, kTC>(ta, B, k0 + kTC, N, size2d, x, xs, m0, mean, p);
    }
    if (k0 < N) {
        thomas_tab_up<kTC>(Z, fs, k0 + kTC, N - 1, B);
        if (k0 + kTC <= N) thomas_fwd_chunk<T, true, kTC>(ta, A, k0, N, size2d, x, xs, m0, mean, p);
        else thomas_fwd_chunk<T, false, kTC>(ta, A, k0, N, size2d, x, xs, m0, mean, p);
        if (k0 + kTC < N) thomas_fwd_chunk<T, false, kTC>(ta, B, k0 + kTC, N, size2d, x, xs, m0, mean, p);
    }
    if (m0) p = 0.0;  // OUT3D(0,0,NZ-2) = 0
    x[(N - 1) * xs] = T(p);
    if (N < 2) return;
    // back substitution x_kk = y_kk - F_kk x_{kk+1} for kk = N-2 .. 0, walking down in chunks
    thomas_tab_down<kTC>(F, fs, N - 2, A);
    k0 = N - 2;
    for (; k0 - 2 * kTC + 1 >= 0; k0 -= 2 * kTC) {
        thomas_tab_down<kTC>(F, fs, k0 - kTC, B);
        thomas_back_chunk<T, true, kTC>(A, k0, x, xs, m0, mean, p);
        thomas_tab_down<kTC>(F, fs, k0 - 2 * kTC, A);
        thomas_back_chunk<T, true, kTC>(B, k0 - kTC, x, xs, m0, mean, p);
    }
    if (k0 >= 0) {
        thomas_tab_down<kTC>(F, fs, k0 - kTC, B);
        if (k0 - kTC + 1 >= 0) thomas_back_chunk<T, true, kTC>(A, k0, x, xs, m0, mean, p);
        else thomas_back_chunk<T, false, kTC>(A, k0, x, xs, m0, mean, p);
        if (k0 - kTC >= 0) thomas_back_chunk<T, false, kTC>(B, k0 - kTC, x, xs, m0, mean, p);
    }
}

// q_kk = x_kk - F_{kk-1} q_{kk-1} for the chunk's levels kk = k0 .. k0+kTC-1 (< N unless FULL), f[t] = F_{k0+t-1}
template <typename T, bool FULL, int kTC>
__device__ __forceinline__ void thomas_t_fwd_chunk(const double (&f)[kTC], int k0, int N, T *x, int xs, double &q) {
    double v[kTC];
#pragma unroll
    for (int t = 0; t < kTC; t++) v[t] = double(x[(FULL ? k0 + t : min(k0 + t, N - 1)) * xs]);
#pragma unroll
    for (int t = 0; t < kTC; t++) {
        if (FULL || k0 + t < N) q = v[t] - f[t] * q;
        v[t] = q;
    }
#pragma unroll
    for (int t = 0; t < kTC; t++)
        if (FULL || k0 + t < N) x[(k0 + t) * xs] = T(v[t]);
}
// b_kk = z_kk (q_kk - a_{kk+2} b_{kk+1}), stored as b / (ni nj), for kk = k0 .. k0-kTC+1 (>= 0 unless FULL)
template <typename T, bool FULL, int kTC>
__device__ __forceinline__ void thomas_t_back_chunk(const ThomasArgs &ta, const double (&z)[kTC], int k0, double size2d, T *x, int xs,
                                                    double &b) {
    double v[kTC], ac[kTC];
#pragma unroll
    for (int t = 0; t < kTC; t++) {
        const int kk = FULL ? k0 - t : max(k0 - t, 0);
        v[t] = double(x[kk * xs]);
        ac[t] = ta.acoef[kk + 2];
    }
#pragma unroll
    for (int t = 0; t < kTC; t++) {
        if (FULL || k0 - t >= 0) b = (v[t] - ac[t] * b) * z[t];
        v[t] = b / size2d;
    }
#pragma unroll
    for (int t = 0; t < kTC; t++)
        if (FULL || k0 - t >= 0) x[(k0 - t) * xs] = T(v[t]);
}
// Transpose of thomas_part (the adjoint's vertical solve), in place.  The forward map is p = L'^{-1} x / S
// (L' lower bidiagonal: 1/z_k on the diagonal, a_k below) followed by x_k = p_k - F_k x_{k+1}; for the
// (0,0) column p_N is replaced by 0.  The transpose runs the two bidiagonal solves transposed and in
// reverse order: q_1 = x_1, q_k = x_k - F_{k-1} q_{k-1} (q_N = 0 for (0,0)), then b_N = z_N q_N,
// b_k = z_k (q_k - a_{k+1} b_{k+1}), result b / S.  The singular last pivot of (0,0) is never used.
// The table entries come in chunks through two register sets that take turns, as in thomas_part.
template <typename T, int kTC = kThomasChunk>
__device__ void thomas_part_t(const ThomasArgs &ta, int kx, int ky, T *x, int xs) {
    const int N = ta.nk - 2;
    const long long fs = ta.nj, col = (long long)kx * N * ta.nj + ky;
    const double *F = ta.ffac + col, *Z = ta.zpiv + col;
    const double size2d = (double)(ta.ni * ta.nj);
    const bool m0 = (kx == 0 && ky == 0);
    double A[kTC], B[kTC];
    // q_kk = x_kk - F_{kk-1} q_{kk-1}, kk = 1 .. N-1 (q_0 = x_0)
    double q = double(x[0]);
    thomas_tab_up<kTC>(F, fs, 0, N - 2, A);
    int k0 = 1;
    for (; k0 + 2 * kTC <= N; k0 += 2 * kTC) {
        thomas_tab_up<kTC>(F, fs, k0 + kTC - 1, N - 2, B);
        thomas_t_fwd_chunk<T, true, kTC>(A, k0, N, x, xs, q);
        thomas_tab_up<kTC>(F, fs, k0 + 2 * kTC - 1, N - 2, A);
        thomas_t_fwd_chunk<T, true, kTC>(B, k0 + kTC, N, x, xs, q);
    }
    if (k0 < N) {
        thomas_tab_up<kTC>(F, fs, k0 + kTC - 1, N - 2, B);
        if (k0 + kTC <= N) thomas_t_fwd_chunk<T, true, kTC>(A, k0, N, x, xs, q);
        else thomas_t_fwd_chunk<T, false, kTC>(A, k0, N, x, xs, q);
        if (k0 + kTC < N) thomas_t_fwd_chunk<T, false, kTC>(B, k0 + kTC, N, x, xs, q);
    }
    // b_{N-1} = z_{N-1} q_{N-1} (0 for (0,0)); b_kk = z_kk (q_kk - a_{kk+2} b_{kk+1}), kk = N-2 .. 0
    double b = m0 ? 0.0 : double(x[(N - 1) * xs]) * Z[(N - 1) * fs];
    x[(N - 1) * xs] = T(b / size2d);
    if (N < 2) return;
    thomas_tab_down<kTC>(Z, fs, N - 2, A);
    k0 = N - 2;
    for (; k0 - 2 * kTC + 1 >= 0; k0 -= 2 * kTC) {
        thomas_tab_down<kTC>(Z, fs, k0 - kTC, B);
        thomas_t_back_chunk<T, true, kTC>(ta, A, k0, size2d, x, xs, b);
        thomas_tab_down<kTC>(Z, fs, k0 - 2 * kTC, A);
        thomas_t_back_chunk<T, true, kTC>(ta, B, k0 - kTC, size2d, x, xs, b);
    }
    if (k0 >= 0) {
        thomas_tab_down<kTC>(Z, fs, k0 - kTC, B);
        if (k0 - kTC + 1 >= 0) thomas_t_back_chunk<T, true, kTC>(ta, A, k0, size2d, x, xs, b);
        else thomas_t_back_chunk<T, false, kTC>(ta, A, k0, size2d, x, xs, b);
        if (k0 - kTC >= 0) thomas_t_back_chunk<T, false, kTC>(ta, B, k0 - kTC, size2d, x, xs, b);
    }
}

// Global <-> LDS copies of one kx slab: n complex values (re, im interleaved) at spec[base + 2 idx], in
// the LDS planes at (idx / nk2) * P2 + idx % nk2.  Each thread issues kU global loads (clamped indices, so
// no load sits behind a branch) before it writes any of them: a guarded load inside the loop made the
// compiler wait for each one before issuing the next.
constexpr int kSlabU = 8;
template <typename T>
__device__ __forceinline__ void slab_to_lds(const T *__restrict__ spec, long long base, int n, int nk2, int P2, T *re, T *im) {
    const int nt = (int)blockDim.x, dk = nt % nk2, dy = nt / nk2;
    int kk = (int)threadIdx.x % nk2, ky = (int)threadIdx.x / nk2;
    for (int i0 = (int)threadIdx.x; i0 < n; i0 += kSlabU * nt) {
        T vr[kSlabU], vi[kSlabU];
#pragma unroll
        for (int u = 0; u < kSlabU; u++) {
            ld2(spec + base + 2LL * min(i0 + u * nt, n - 1), vr[u], vi[u]);
        }
#pragma unroll
        for (int u = 0; u < kSlabU; u++) {
            if (i0 + u * nt < n) {
                re[ky * P2 + kk] = vr[u];
                im[ky * P2 + kk] = vi[u];
            }
            kk += dk;
            ky += dy;
            if (kk >= nk2) {
                kk -= nk2;
                ky++;
            }
        }
    }
}
template <typename T>
__device__ __forceinline__ void lds_to_slab(T *__restrict__ spec, long long base, int n, int nk2, int P2, const T *re, const T *im) {
    const int nt = (int)blockDim.x, dk = nt % nk2, dy = nt / nk2, last = n / nk2 - 1;
    int kk = (int)threadIdx.x % nk2, ky = (int)threadIdx.x / nk2;
    for (int i0 = (int)threadIdx.x; i0 < n; i0 += kSlabU * nt) {
        T vr[kSlabU], vi[kSlabU];
#pragma unroll
        for (int u = 0; u < kSlabU; u++) {
            const int e = min(ky, last) * P2 + kk;  // rows past the slab read its last row (not stored)
            vr[u] = re[e];
            vi[u] = im[e];
            kk += dk;
            ky += dy;
            if (kk >= nk2) {
                kk -= nk2;
                ky++;
            }
        }
#pragma unroll
        for (int u = 0; u < kSlabU; u++) {
            const int idx = i0 + u * nt;
            if (idx < n) st2(spec + base + 2LL * idx, vr[u], vi[u]);
        }
    }
}

// Q2 (fused): forward DFT along j, FP64 tridiagonal solve, backward DFT along j for one kx, with the
// whole (j, k) slab of that kx resident in LDS
template <typename T, bool TR, bool POW2 = true>
__device__ __forceinline__ void ypass_block(T *__restrict__ spec, int nk2, const Radix &rady, const double *twy, const ThomasArgs &ta,
                                            int pad, int kx, unsigned char *smem) {
    const int nj = rady.n, P2 = nk2 + pad;
    T *re = (T *)smem, *im = re + nj * P2;
    double *mean = (double *)(re + 2 * nj * P2);  // 2*nj*P2 elements of T: 8-byte aligned
    T *twl = (T *)(mean + 2 * nk2 + 1);
    load_twiddles(twl, twy, nj);
    const long long base = (long long)kx * nj * nk2 * 2;
    slab_to_lds(spec, base, nj * nk2, nk2, P2, re, im);
    __syncthreads();
    lds_fft_ip<T, ypass_maxb<T>(), POW2>(re, im, rady, nk2, P2, 1, twl, -1);
    if ((int)threadIdx.x < 2 * nj) {
        // one thread per real sequence: threads [0, nj) take the real parts, [nj, 2 nj) the imaginary parts
        const int part = (int)threadIdx.x >= nj, ky = (int)threadIdx.x - part * nj;
        T *x = (part ? im : re) + ky * P2;
        if (TR) thomas_part_t<T>(ta, kx, ky, x, 1);
        else thomas_part<T>(ta, kx, ky, x, 1, mean + part);
    }
    __syncthreads();
    lds_fft_ip<T, ypass_maxb<T>(), POW2>(re, im, rady, nk2, P2, 1, twl, +1);
    lds_to_slab(spec, base, nj * nk2, nk2, P2, re, im);
}

template <typename T>
__global__ void __launch_bounds__(kThreadsX) k_div_rfft_x(Geo g, Cols c, const T *__restrict__ UP, const T *__restrict__ VP,
                                                         const T *__restrict__ WP, T *__restrict__ spec, Radix rad, const double *tw,
                                                         int nk2, int kc, double step, const T *__restrict__ src) {
    extern __shared__ unsigned char smem[];
    div_rfft_x_block<T>(g, c, UP, VP, WP, spec, rad, tw, nk2, kc, step, src, blockIdx.x, blockIdx.y, smem);
}

template <typename T>
__global__ void __launch_bounds__(kThreadsX) k_irfft_x_out(Geo g, const T *__restrict__ spec, T *__restrict__ PI, Radix rad,
                                                          const double *tw, int nk2, int kc, int half) {
    extern __shared__ unsigned char smem[];
    irfft_x_out_block<T>(g, spec, PI, rad, tw, nk2, kc, blockIdx.x, blockIdx.y, smem, half);
}

// The same after the half-slab j pass, one block per row pair (j', jhalf + j') and chunk of kc levels:
// it reads v0 (row j') and v1 (row jhalf + j') once for both rows (v0 + v1 and v0 - v1), and transforms
// the two rows as one batch of kc level pairs.  Values as irfft_x_out_block's, bit for bit.
template <typename T>
__device__ __forceinline__ void irfft_x_out_pair_block(const Geo &g, const T *__restrict__ spec, T *__restrict__ PI, const Radix &rad,
                                                       const double *tw, int nk2, int kc, int jp, int chunk, unsigned char *smem,
                                                       int jhalf) {
    const int n = rad.n, P = kc / 2, B = 2 * P, half = n / 2;  // B: batch of both rows' level pairs
    T *re = (T *)smem, *im = re + n * B, *twl = im + n * B;
    const int kk0 = chunk * kc;
    load_twiddles(twl, tw, n);
#pragma unroll 2
    for (int idx = threadIdx.x; idx < n * P; idx += blockDim.x) {
        int pp = idx % P, m = idx / P;
        int ka = kk0 + 2 * pp, kb = ka + 1;
        int ms = m <= half ? m : n - m;
        const long long o0 = ((long long)ms * g.NY + jp) * nk2, o1 = ((long long)ms * g.NY + jhalf + jp) * nk2;
        T e[4] = {T(0), T(0), T(0), T(0)}, o[4] = {T(0), T(0), T(0), T(0)};  // a re, a im, b re, b im
        if (kb < nk2 && (nk2 & 1) == 0) {  // the pair's four values in one load
            ld4(spec + (o0 + ka) * 2, e);
            ld4(spec + (o1 + ka) * 2, o);
        } else {
            if (ka < nk2) {
                ld2(spec + (o0 + ka) * 2, e[0], e[1]);
                ld2(spec + (o1 + ka) * 2, o[0], o[1]);
            }
            if (kb < nk2) {
                ld2(spec + (o0 + kb) * 2, e[2], e[3]);
                ld2(spec + (o1 + kb) * 2, o[2], o[3]);
            }
        }
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const T sg = r == 0 ? T(1) : T(-1);
            T ar = T(0), ai = T(0), br = T(0), bi = T(0);
            if (ka < nk2) {
                ar = e[0] + sg * o[0];
                ai = e[1] + sg * o[1];
            }
            if (kb < nk2) {
                br = e[2] + sg * o[2];
                bi = e[3] + sg * o[3];
            }
            if (ms == 0 || ms == half) {  // self-conjugate wavenumbers: only the real part is kept
                ai = T(0);
                bi = T(0);
            }
            if (m > half) {  // conjugate mirror
                ai = -ai;
                bi = -bi;
            }
            re[m * B + r * P + pp] = ar - bi;
            im[m * B + r * P + pp] = ai + br;
        }
    }
    __syncthreads();
    lds_fft_ip<T, kMaxBX>(re, im, rad, B, B, 1, twl, +1);
    const int shift = g.NX - 6;
    for (int idx = threadIdx.x; idx < 2 * n * kc; idx += blockDim.x) {
        const int r = idx / (n * kc), rem = idx - r * (n * kc);
        const int b = rem % kc, i = rem / kc;
        const int kk = kk0 + b;
        if (kk >= nk2) continue;
        const int q = r * P + (b >> 1);
        T v = (b & 1) ? im[i * B + q] : re[i * B + q];
        const int j = r == 0 ? jp : jhalf + jp;
        int ia = i + 3, k = kk + 1;
        int ks[3] = {k, -1, -1};
        if (kk == 0) ks[1] = 0;
        if (kk == nk2 - 1) ks[2] = g.NZ - 1;
        for (int t = 0; t < 3; t++) {
            if (ks[t] < 0) continue;
            PI[IX(ia, j, ks[t])] = v;
            if (ia >= shift && ia < shift + 3) PI[IX(ia - shift, j, ks[t])] = v;
            if (ia >= 3 && ia < 6) PI[IX(ia + shift, j, ks[t])] = v;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreadsX) k_irfft_x_out_pair(Geo g, const T *__restrict__ spec, T *__restrict__ PI, Radix rad,
                                                               const double *tw, int nk2, int kc, int jhalf) {
    extern __shared__ unsigned char smem[];
    irfft_x_out_pair_block<T>(g, spec, PI, rad, tw, nk2, kc, blockIdx.x, blockIdx.y, smem, jhalf);
}

// Q2 on half slabs (even nj, L = nj/2): the j transform's first radix-2 split (decimation in frequency)
// is done while loading, so block (kx, h) holds only a_h(j') = (x(j') + (-1)^h x(j'+L)) W^(h j'),
// W = exp(-2 pi i / nj), j' < L, whose length-L transform is X(2m + h).  It solves the columns
// ky = 2m + h, transforms back, and stores v_h(j') = conj(W)^(h j') u_h(j') in rows h L + j' of `out`; the
// inverse along i combines row j of the full inverse as v_0(j') + (-1)^q v_1(j'), j = j' + qL.  Both
// blocks of a kx read every row of `in`, so they write to a second array.
template <typename T, bool TR>
__device__ __forceinline__ void ypass_half_block(const T *__restrict__ in, T *__restrict__ out, int nk2, const Radix &rh,
                                                 const double *twy, const ThomasArgs &ta, int pad, int kx, int h, unsigned char *smem) {
    const int L = rh.n, nj = 2 * L, P2 = nk2 + pad;
    T *re = (T *)smem, *im = re + L * P2;
    double *mean = (double *)(re + 2 * L * P2);
    T *twl = (T *)(mean + 2 * nk2 + 1);
    // the length-L table is every other entry of the length-nj one
    for (int e = threadIdx.x; e < L; e += blockDim.x) {
        twl[2 * e] = T(twy[4 * e]);
        twl[2 * e + 1] = T(twy[4 * e + 1]);
    }
    const long long base = (long long)kx * nj * nk2 * 2, second = 2LL * L * nk2;
    // kU elements per thread per round, every global load of a round issued (clamped indices) before any
    // is used: a load behind the loop's bound check made the compiler wait for each in turn
    constexpr int kU = 4;
    const int ne = L * nk2, nt = (int)blockDim.x;
    for (int idx0 = (int)threadIdx.x; idx0 < ne; idx0 += kU * nt) {
        T a[kU][4];
        double w[kU][2];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int idx = min(idx0 + u * nt, ne - 1), jp = idx / nk2;
            const long long o = base + 2LL * idx;
            ld2(in + o, a[u][0], a[u][1]);
            ld2(in + o + second, a[u][2], a[u][3]);
            if (h == 1) {
                w[u][0] = twy[2 * jp];
                w[u][1] = twy[2 * jp + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int idx = idx0 + u * nt, jp = idx / nk2, kk = idx - jp * nk2;
            T ar, ai;
            if (h == 0) {
                ar = a[u][0] + a[u][2];
                ai = a[u][1] + a[u][3];
            } else {
                const T dr = a[u][0] - a[u][2], di = a[u][1] - a[u][3];
                ar = dr;
                ai = di;
                cmul(ar, ai, T(w[u][0]), T(w[u][1]));
            }
            if (idx < ne) {
                re[jp * P2 + kk] = ar;
                im[jp * P2 + kk] = ai;
            }
        }
    }
    __syncthreads();
    lds_fft_pow2<T, half_maxb<T>()>(re, im, rh, nk2, P2, 1, twl, -1);
    if ((int)threadIdx.x < 2 * L) {
        const int part = (int)threadIdx.x >= L, m = (int)threadIdx.x - part * L;
        T *x = (part ? im : re) + m * P2;
        if (TR) thomas_part_t<T>(ta, kx, 2 * m + h, x, 1);
        else thomas_part<T>(ta, kx, 2 * m + h, x, 1, mean + part);
    }
    __syncthreads();
    lds_fft_pow2<T, half_maxb<T>()>(re, im, rh, nk2, P2, 1, twl, +1);
    for (int idx0 = (int)threadIdx.x; idx0 < ne; idx0 += kU * nt) {
        T vr[kU], vi[kU];
        double w[kU][2];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int idx = min(idx0 + u * nt, ne - 1), jp = idx / nk2, kk = idx - jp * nk2;
            vr[u] = re[jp * P2 + kk];
            vi[u] = im[jp * P2 + kk];
            if (h == 1) {
                w[u][0] = twy[2 * jp];
                w[u][1] = twy[2 * jp + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int idx = idx0 + u * nt;
            if (h == 1) cmul(vr[u], vi[u], T(w[u][0]), -T(w[u][1]));
            if (idx < ne) {
                st2(out + base + 2LL * ((long long)h * L * nk2 + idx), vr[u], vi[u]);
            }
        }
    }
}

template <typename T, bool TR = false>
__global__ void __launch_bounds__(half_threads<T>()) k_ypass_half(const T *__restrict__ in, T *__restrict__ out, int nk2, Radix rh,
                                                                   const double *twy, ThomasArgs ta, int pad) {
    extern __shared__ unsigned char smem[];
    ypass_half_block<T, TR>(in, out, nk2, rh, twy, ta, pad, blockIdx.x >> 1, blockIdx.x & 1, smem);
}

template <typename T, bool TR = false>
__global__ void __launch_bounds__(ypass_threads<T>()) k_ypass(T *__restrict__ spec, int nk2, Radix rady, const double *twy, ThomasArgs ta,
                                                              int pad) {
    extern __shared__ unsigned char smem[];
    ypass_block<T, TR>(spec, nk2, rady, twy, ta, pad, blockIdx.x, smem);
}

// ---- batched over the members of an ensemble: one launch per pass, blocks walk a work list that
// concatenates the members' (row, level chunk) or (kx) items -------------------------------------------

template <typename T>
__global__ void __launch_bounds__(kThreadsX) k_ens_div_rfft_x(const EnsMember *mem, int nmem, EnsRoles roles, int total, double step) {
    extern __shared__ unsigned char smem[];
    {
        const int item = blockIdx.x;
        const EnsMember &e = mem[ens_member(mem, nmem, item, &EnsMember::px0)];
        const int local = item - e.px0, nch = (e.nk2 + e.kcx - 1) / e.kcx;
        div_rfft_x_block<T>(e.g, e.c, (const T *)e.buf[DYN_U][roles.r[DYN_U][DYN_ROLE_P]], (const T *)e.buf[DYN_V][roles.r[DYN_V][DYN_ROLE_P]],
                            (const T *)e.buf[DYN_W][roles.r[DYN_W][DYN_ROLE_P]], (T *)e.spec, e.rx, e.twx, e.nk2, e.kcx, step,
                            (const T *)nullptr, local / nch, local % nch, smem);
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreadsX) k_ens_irfft_x_out(const EnsMember *mem, int nmem, int total) {
    extern __shared__ unsigned char smem[];
    {
        const int item = blockIdx.x;
        const EnsMember &e = mem[ens_member(mem, nmem, item, &EnsMember::px0)];
        const int local = item - e.px0, nch = (e.nk2 + e.kcx - 1) / e.kcx;
        irfft_x_out_block<T>(e.g, (const T *)e.spec, (T *)e.pi, e.rx, e.twx, e.nk2, e.kcx, local / nch, local % nch, smem);
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_ens_ypass(const EnsMember *mem, int nmem, int total) {
    extern __shared__ unsigned char smem[];
    {
        const int item = blockIdx.x;
        const EnsMember &e = mem[ens_member(mem, nmem, item, &EnsMember::py0)];
        ThomasArgs ta{e.coskx, e.cosly, e.c.rhou, e.c.rhow, e.c.mu, e.c.mw, e.acoef, e.ffac, e.dx, e.dz, e.ni, e.nj, e.g.NZ, e.zpiv};
        // the generic transform passes (the straight-line power-of-two ones measured no faster here)
        ypass_block<T, false, false>((T *)e.spec, e.nk2, e.ry, e.twy, ta, 1, item - e.py0, smem);
    }
}

// Q2 (staged): tridiagonal solve of tg columns of one kx staged through LDS (coalesced in and out)
template <typename T, bool TR = false>
__global__ void __launch_bounds__(128) k_thomas_staged(T *__restrict__ spec, int nk2, ThomasArgs ta) {
    extern __shared__ unsigned char smem[];
    const int tg = blockDim.x / 2, P2 = 2 * nk2 + 1;
    T *x = (T *)smem;
    double *mean = (double *)(smem + ((tg * P2 * sizeof(T) + 7) / 8) * 8);
    const int kx = blockIdx.x, ky0 = blockIdx.y * tg;
    const int ncol = min(tg, ta.nj - ky0);
    const long long base = ((long long)kx * ta.nj + ky0) * nk2 * 2;
    for (int idx = threadIdx.x; idx < ncol * nk2 * 2; idx += blockDim.x) {
        int col = idx / (2 * nk2), rem = idx % (2 * nk2);
        x[col * P2 + rem] = spec[base + idx];
    }
    __syncthreads();
    {
        // two threads per column: even lanes the real part, odd lanes the imaginary part (interleaved)
        const int colx = (int)threadIdx.x >> 1, part = threadIdx.x & 1;
        if (colx < ncol) {
            if (TR) thomas_part_t<T>(ta, kx, ky0 + colx, x + colx * P2 + part, 2);
            else thomas_part<T>(ta, kx, ky0 + colx, x + colx * P2 + part, 2, mean + part);
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < ncol * nk2 * 2; idx += blockDim.x) {
        int col = idx / (2 * nk2), rem = idx % (2 * nk2);
        spec[base + idx] = x[col * P2 + rem];
    }
}

// Radix passes of a length-n transform: 4, 2, 3 and 5 first (the butterflies), then the remaining prime
// factors in increasing order (direct sums, dft_radix); false when n needs more than kMaxRad passes.
bool factorize(int n, Radix &rad) {
    rad.n = n;
    rad.nr = 0;
    int m = n;
    while (m > 1 && rad.nr < kMaxRad) {
        int r = (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : 0;
        if (r == 0) {
            r = 7;
            while (m % r != 0) r++;  // smallest prime factor of m (every smaller factor was divided out)
        }
        rad.r[rad.nr++] = r;
        m /= r;
    }
    return m == 1;
}

int level_chunk(int n, size_t elt, int nk2);
size_t x_pass_lds(int ni, int kc, size_t elt);

// the inverse i pass into `out`: after the half-slab j pass (half = nj/2) one block per row pair and
// chunk of an even number of levels, at most kc/2 (same LDS and butterflies per thread as kc levels of
// one row); otherwise one block per row and chunk of kc levels
template <typename T>
void launch_irfft_x_out(Ctx *c, const T *spec, T *out, const Radix &rx, int kc, int half) {
    const PressurePlan &pp = c->pp;
    const Geo &g = c->geo;
    const size_t shx = x_pass_lds(pp.ni, kc, sizeof(T));
    const int kcp = (kc / 2) & ~1;
    if (half && kcp >= 2) {
        hipLaunchKernelGGL(k_irfft_x_out_pair<T>, dim3(half, (pp.nk2 + kcp - 1) / kcp), dim3(kThreadsX), shx, c->stream, g, spec, out, rx,
                           pp.twx, pp.nk2, kcp, half);
        return;
    }
    hipLaunchKernelGGL(k_irfft_x_out<T>, dim3(g.NY, (pp.nk2 + kc - 1) / kc), dim3(kThreadsX), shx, c->stream, g, spec, out, rx, pp.twx,
                       pp.nk2, kc, half);
}

// one j transform of the three-kernel j pass over nrows kx rows: the power-of-two kernel when nj is a
// power of two and its plan fits, else the mixed-radix one
template <typename T>
void launch_fft_y(Ctx *c, T *spec, int nrows, const Radix &ry, int sign) {
    const PressurePlan &pp = c->pp;
    Radix rp{};
    const int kcp = pow2_plan48(pp.nj, rp, true) ? fft_y_pow2_chunk<T>(rp, pp.nk2) : 0;
    if (kcp > 0) {
        hipLaunchKernelGGL((k_fft_y_pow2<T, kFftYMaxB>), dim3(nrows, (pp.nk2 + kcp - 1) / kcp), dim3(kFftYThreads),
                           fft_y_pow2_lds(pp.nj, kcp, sizeof(T)), c->stream, spec, pp.nj, pp.nk2, rp, pp.twy, kcp, sign);
        return;
    }
    const int kcy = level_chunk(pp.nj, sizeof(T), pp.nk2);
    hipLaunchKernelGGL(k_fft_y<T>, dim3(nrows, (pp.nk2 + kcy - 1) / kcy), dim3(kThreads), 4 * (size_t)pp.nj * kcy * sizeof(T), c->stream,
                       spec, pp.nj, pp.nk2, ry, pp.twy, kcy, sign);
}

int level_chunk(int n, size_t elt, int nk2) {
    // two ping-pong buffers of n*kc complex values must fit the LDS budget
    int kc = (int)(kLdsBudget / (4 * (size_t)n * elt));
    if (kc > 32) kc = 32;
    if (kc > nk2) kc = nk2;
    if (kc < 1) kc = 1;
    return kc;
}

double *upload_twiddles(int n, hipStream_t s) {
    std::vector<double> h(2 * (size_t)n);
    for (int e = 0; e < n; e++) {
        double a = 2.0 * kTrigPi * (double)e / (double)n;
        h[2 * e] = std::cos(a);
        h[2 * e + 1] = -std::sin(a);
    }
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

// butterflies a thread holds in the worst pass; a pass of prime radix above 5 holds outputs instead, 5
// per radix-5 butterfly's worth of registers
int max_butterflies(const Radix &r, int nbatch, int nthr) {
    int mb = 0;
    for (int s = 0; s < r.nr; s++) {
        if (r.r[s] > 5) mb = std::max(mb, ((r.n * nbatch + nthr - 1) / nthr + 4) / 5);
        else mb = std::max(mb, (r.n / r.r[s] * nbatch + nthr - 1) / nthr);
    }
    return mb;
}

// LDS of the i passes for kc levels (kc/2 packed pairs) and of the fused j pass, twiddle tables included
size_t x_pass_lds(int ni, int kc, size_t elt) { return (2 * (size_t)ni * (kc / 2) + 2 * (size_t)ni) * elt + 32 * (size_t)kc; }
size_t y_pass_lds(int nj, int nk2, size_t elt) {
    return 2 * (size_t)nj * (nk2 + 1) * elt + 8 * (2 * (size_t)nk2 + 1) + 2 * (size_t)nj * elt + 8;
}
// levels per block of the i passes (even: two levels per complex transform), within the LDS budget and
// the butterflies a thread holds.  One grid alone wants enough blocks to fill the chip (about 512 rows x
// chunks); a batch of members has blocks to spare, so there every row takes all its levels at once
// (`fill` false), which keeps the row segments long.
int x_level_chunk(const PressurePlan &pp, const Geo &g, const Radix &rx, size_t elt, bool fill = true) {
    int kc = (int)((kLdsBudget - 2 * (size_t)pp.ni * elt) / (2 * (size_t)pp.ni * elt)) * 2;
    // about 512 blocks per pass in FP32; FP64 takes half the levels per block down to 16 (256x256x64:
    // 1.909 -> 1.869 ms per step with 16 levels instead of 30; FP32 is slower with fewer than 32, and
    // FP64 at 192x192x40 with 6 instead of 14)
    int want = fill ? std::max(2, ((pp.nk2 * g.NY / 512) + 1) & ~1) : (pp.nk2 + 1) & ~1;
    if (fill && elt == 8 && want > 16) want = std::max(16, (want / 2) & ~1);
    kc = std::max(2, std::min(kc, want));
    while (kc > 2 && (max_butterflies(rx, kc / 2, kThreadsX) > kMaxBX || x_pass_lds(pp.ni, kc, elt) > kLdsBudget)) kc -= 2;
    return kc;
}

template <typename T>
int solve_t(Ctx *c, double step, bool halos, PressureHook after_i, void *arg) {
    PressurePlan &pp = c->pp;
    const Geo &g = c->geo;
    Radix rx{}, ry{};
    factorize(pp.ni, rx);
    factorize(pp.nj, ry);
    if (halos) {  // periodic_uvw_boundaries (boundaries.cpp:546-555)
        void *uvw[3] = {c->prog[DYN_U][2], c->prog[DYN_V][2], c->prog[DYN_W][2]};
        int rc = launch_periodic(c, uvw, 3);
        if (rc) return rc;
    }
    T *spec = (T *)pp.spec;
    const double pts = (double)pp.ni * pp.nj * pp.nk2;
    ThomasArgs ta{pp.coskx, pp.cosly, c->cols.rhou, c->cols.rhow, c->cols.mu, c->cols.mw, pp.acoef, pp.ffac, c->g.dx, c->g.dz,
                  pp.ni, pp.nj, c->g.nz, pp.zpiv};
    // even ni of factors 2, 3, 5: half-spectrum transforms in place; otherwise the full complex spectrum
    if (pp.ni % 2 == 0 && pp.smooth_x) {
        const int nkx = pp.ni / 2 + 1;
        const int kc = x_level_chunk(pp, g, rx, sizeof(T));
        const size_t shx = x_pass_lds(pp.ni, kc, sizeof(T));
        dim3 gx(g.NY, (pp.nk2 + kc - 1) / kc);
        {
            Timed _t(c, "p1_div_rfft_x", pts);
            hipLaunchKernelGGL(k_div_rfft_x<T>, gx, dim3(kThreadsX), shx, c->stream, g, c->cols, (const T *)c->prog[DYN_U][2],
                               (const T *)c->prog[DYN_V][2], (const T *)c->prog[DYN_W][2], spec, rx, pp.twx, pp.nk2, kc, step,
                               (const T *)nullptr);
        }
        DYN_LAUNCH_CHECK();
        if (after_i) {
            const int rc = after_i(c, arg);
            if (rc) return rc;
        }
        const int pad = 1;
        const size_t shy = y_pass_lds(pp.nj, pp.nk2, sizeof(T));
        const int jpass = fused_shape(c, DYN_TUNE_PRESSURE_JPASS);
        int half = 0;
        if (jpass == 1) {
            Timed _t(c, "p2_ypass", pts);
            hipLaunchKernelGGL(k_ypass<T>, dim3(nkx), dim3(ypass_threads<T>()), shy, c->stream, spec, pp.nk2, ry, pp.twy, ta, pad);
            DYN_LAUNCH_CHECK();
        } else if (jpass == 3) {
            Radix rh{};
            pow2_plan48(pp.nj / 2, rh, sizeof(T) == 8);
            half = pp.nj / 2;
            Timed _t(c, "p2_ypass", pts);
            hipLaunchKernelGGL((k_ypass_half<T>), dim3(2 * nkx), dim3(half_threads<T>()), y_pass_lds(half, pp.nk2, sizeof(T)), c->stream,
                               (const T *)spec, (T *)pp.spec2, pp.nk2, rh, pp.twy, ta, pad);
            DYN_LAUNCH_CHECK();
        } else {
            {
                Timed _t(c, "p2_fft_y", pts);
                launch_fft_y<T>(c, spec, nkx, ry, -1);
            }
            DYN_LAUNCH_CHECK();
            const int tg = 64;
            size_t sht = ((tg * (2 * (size_t)pp.nk2 + 1) * sizeof(T) + 7) / 8) * 8 + 16 * (size_t)pp.nk2;
            {
                Timed _t(c, "p3_thomas", pts);
                hipLaunchKernelGGL(k_thomas_staged<T>, dim3(nkx, (pp.nj + tg - 1) / tg), dim3(2 * tg), sht, c->stream, spec, pp.nk2, ta);
            }
            DYN_LAUNCH_CHECK();
            {
                Timed _t(c, "p4_ifft_y", pts);
                launch_fft_y<T>(c, spec, nkx, ry, +1);
            }
            DYN_LAUNCH_CHECK();
        }
        {
            Timed _t(c, "p5_irfft_x_out", pts);
            launch_irfft_x_out<T>(c, half ? (const T *)pp.spec2 : spec, (T *)c->pi, rx, kc, half);
        }
        DYN_LAUNCH_CHECK();
        return DYN_OK;
    }
    // odd ni, or a prime factor above 5: full complex spectrum, out-of-place transforms
    int kcx = level_chunk(pp.ni, sizeof(T), pp.nk2), kcy = level_chunk(pp.nj, sizeof(T), pp.nk2);
    size_t shx = 4 * (size_t)pp.ni * kcx * sizeof(T), shy = 4 * (size_t)pp.nj * kcy * sizeof(T);
    dim3 gx(g.NY, (pp.nk2 + kcx - 1) / kcx), gy(pp.ni, (pp.nk2 + kcy - 1) / kcy);
    {
        Timed _t(c, "p1_div_fft_x", pts);
        hipLaunchKernelGGL(k_div_fft_x<T>, gx, dim3(kThreads), shx, c->stream, g, c->cols, (const T *)c->prog[DYN_U][2],
                           (const T *)c->prog[DYN_V][2], (const T *)c->prog[DYN_W][2], spec, rx, pp.twx, pp.nk2, kcx, step);
    }
    DYN_LAUNCH_CHECK();
    if (after_i) {
        const int rc = after_i(c, arg);
        if (rc) return rc;
    }
    {
        Timed _t(c, "p2_fft_y", pts);
        hipLaunchKernelGGL(k_fft_y<T>, gy, dim3(kThreads), shy, c->stream, spec, pp.nj, pp.nk2, ry, pp.twy, kcy, -1);
    }
    DYN_LAUNCH_CHECK();
    int nq = pp.ni * pp.nj;
    {
        Timed _t(c, "p3_thomas", pts);
        hipLaunchKernelGGL(k_thomas<T>, dim3((nq + 127) / 128), dim3(128), 0, c->stream, spec, pp.ni, pp.nj, g.NZ, pp.coskx, pp.cosly,
                           c->cols.rhou, c->cols.rhow, c->cols.mu, c->cols.mw, pp.acoef, pp.ffac, c->g.dx, c->g.dz);
    }
    DYN_LAUNCH_CHECK();
    {
        Timed _t(c, "p4_ifft_y", pts);
        hipLaunchKernelGGL(k_fft_y<T>, gy, dim3(kThreads), shy, c->stream, spec, pp.nj, pp.nk2, ry, pp.twy, kcy, +1);
    }
    DYN_LAUNCH_CHECK();
    {
        Timed _t(c, "p5_ifft_x_out", pts);
        hipLaunchKernelGGL(k_ifft_x_out<T>, gx, dim3(kThreads), shx, c->stream, g, (const T *)spec, (T *)c->pi, rx, pp.twx, pp.nk2, kcx);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}


// ---- adjoint of the solve (transposed projection for the hand-written adjoint) -----------------------
// Transpose of P5's output stage: the value v(i, j, kk) was written to PI at (i+3, j, kk+1), at the ghost
// level 0 (kk = 0) or NZ-1 (kk = nk2-1), and at the periodic halo copies; its adjoint is the sum of the
// adjoints of every location it was written to.
template <typename T>
__global__ void k_adj_pi_fold(Geo g, const T *__restrict__ PH, T *__restrict__ R, int nk2) {
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, ia = blockIdx.z + 3;
    if (kk >= nk2) return;
    const int shift = g.NX - 6, k = kk + 1;
    const int ks[3] = {k, kk == 0 ? 0 : -1, kk == nk2 - 1 ? g.NZ - 1 : -1};
    T v = T(0);
    for (int t = 0; t < 3; t++) {
        if (ks[t] < 0) continue;
        v += PH[IX(ia, j, ks[t])];
        if (ia >= shift && ia < shift + 3) v += PH[IX(ia - shift, j, ks[t])];
        if (ia >= 3 && ia < 6) v += PH[IX(ia + shift, j, ks[t])];
    }
    R[IX(ia, j, k)] = v;
}

// Transpose of P1's divergence prologue (divergence3d_periodic, pressure.cpp:1390-1424): the adjoint of the
// right-hand side r(ia, j, k) (rows 0 and NY-1 fold onto the mirrored rows 1 and NY-2) is spread back onto
// the forecast velocities, one gathering thread per velocity point.
template <typename T>
__global__ void k_adj_div(Geo g, Cols c, const T *__restrict__ R, T *__restrict__ UH, T *__restrict__ VH, T *__restrict__ WH,
                          double step) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)g.NX * g.NYNZ) return;
    const int x = (int)(q / g.NYNZ), r = (int)(q - (long long)x * g.NYNZ), y = r / g.NZ, k = r - y * g.NZ;
    const T sdt = T(step * g.dt);
    auto Rf = [&](int ia, int yy, int kk) -> T {
        if (ia < 3 || ia >= g.NX - 3 || yy < 1 || yy > g.NY - 2 || kk < 1 || kk > g.NZ - 2) return T(0);
        T v = R[IX(ia, yy, kk)];
        if (yy == 1) v += R[IX(ia, 0, kk)];
        if (yy == g.NY - 2) v += R[IX(ia, g.NY - 1, kk)];
        return v;
    };
    T uh = T(0), vh = T(0), wh = T(0);
    if (x >= 3 && x < g.NX - 3 && k >= 1 && k <= g.NZ - 2) {
        const T a = T(c.rhou[k]) * T(g.one_d_dx) / sdt, b = T(c.rhou[k]) * T(g.one_d_dy) / sdt;
        // UP(x) enters column ia as UP(ie) with ie = ia + 1 (ie = 3 for ia = NX-4) and as UP(ia)
        const int iae = x == 3 ? g.NX - 4 : x - 1;
        uh = a * (Rf(iae, y, k) - Rf(x, y, k));
        vh = b * (Rf(x, y - 1, k) - Rf(x, y, k));
    }
    if (x >= 3 && x < g.NX - 3) {
        const T rw = T(c.rhow[k]), odz = T(g.one_d_dz);
        if (k >= 2) wh += rw * (odz * T(c.mu[k - 1])) * Rf(x, y, k - 1) / sdt;
        if (k >= 1 && k <= g.NZ - 2) wh -= rw * (odz * T(c.mu[k])) * Rf(x, y, k) / sdt;
    }
    const long long p = IX(x, y, k);
    UH[p] += uh;
    VH[p] += vh;
    WH[p] += wh;
}

template <typename T>
__global__ void k_zero_field(T *__restrict__ A, long long n) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) A[p] = T(0);
}

template <typename T>
int solve_adjoint_t(Ctx *c, double step, T *pihat, T *uh, T *vh, T *wh, T *r1, T *r2) {
    PressurePlan &pp = c->pp;
    const Geo &g = c->geo;
    if (pp.ni % 2 != 0 || !pp.smooth_x) {
        set_error("adjoint pressure solve: needs an even ni of factors 2, 3 and 5 (half-spectrum transform), got ni=" +
                  std::to_string(pp.ni));
        return DYN_E_UNSUPPORTED;
    }
    Radix rx{}, ry{};
    factorize(pp.ni, rx);
    factorize(pp.nj, ry);
    const double pts = (double)pp.ni * pp.nj * pp.nk2;
    {
        Timed _t(c, "adj_pi_fold", pts);
        hipLaunchKernelGGL(k_adj_pi_fold<T>, dim3((pp.nk2 + 63) / 64, g.NY, pp.ni), dim3(64), 0, c->stream, g, (const T *)pihat, r1, pp.nk2);
        DYN_LAUNCH_CHECK();
        long long n = (long long)c->npts;
        hipLaunchKernelGGL(k_zero_field<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, pihat, n);
        DYN_LAUNCH_CHECK();
    }
    T *spec = (T *)pp.spec;
    ThomasArgs ta{pp.coskx, pp.cosly, c->cols.rhou, c->cols.rhow, c->cols.mu, c->cols.mw, pp.acoef, pp.ffac, c->g.dx, c->g.dz,
                  pp.ni, pp.nj, c->g.nz, pp.zpiv};
    const int nkx = pp.ni / 2 + 1;
    const int kc = x_level_chunk(pp, g, rx, sizeof(T));
    const size_t shx = x_pass_lds(pp.ni, kc, sizeof(T));
    dim3 gx(g.NY, (pp.nk2 + kc - 1) / kc);
    {
        Timed _t(c, "adj_rfft_x", pts);
        hipLaunchKernelGGL(k_div_rfft_x<T>, gx, dim3(kThreadsX), shx, c->stream, g, c->cols, (const T *)nullptr, (const T *)nullptr,
                           (const T *)nullptr, spec, rx, pp.twx, pp.nk2, kc, step, (const T *)r1);
    }
    DYN_LAUNCH_CHECK();
    const int pad = 1;
    const size_t shy = y_pass_lds(pp.nj, pp.nk2, sizeof(T));
    const int jpass = fused_shape(c, DYN_TUNE_PRESSURE_JPASS);
    int half = 0;
    if (jpass == 1) {
        Timed _t(c, "adj_ypass", pts);
        hipLaunchKernelGGL((k_ypass<T, true>), dim3(nkx), dim3(ypass_threads<T>()), shy, c->stream, spec, pp.nk2, ry, pp.twy, ta, pad);
        DYN_LAUNCH_CHECK();
    } else if (jpass == 3) {
        Radix rh{};
        pow2_plan48(pp.nj / 2, rh, sizeof(T) == 8);
        half = pp.nj / 2;
        Timed _t(c, "adj_ypass", pts);
        hipLaunchKernelGGL((k_ypass_half<T, true>), dim3(2 * nkx), dim3(half_threads<T>()), y_pass_lds(half, pp.nk2, sizeof(T)),
                           c->stream, (const T *)spec, (T *)pp.spec2, pp.nk2, rh, pp.twy, ta, pad);
        DYN_LAUNCH_CHECK();
    } else {
        Timed _t(c, "adj_ypass", pts);
        launch_fft_y<T>(c, spec, nkx, ry, -1);
        DYN_LAUNCH_CHECK();
        const int tg = 64;
        size_t sht = ((tg * (2 * (size_t)pp.nk2 + 1) * sizeof(T) + 7) / 8) * 8 + 16 * (size_t)pp.nk2;
        hipLaunchKernelGGL((k_thomas_staged<T, true>), dim3(nkx, (pp.nj + tg - 1) / tg), dim3(2 * tg), sht, c->stream, spec, pp.nk2, ta);
        DYN_LAUNCH_CHECK();
        launch_fft_y<T>(c, spec, nkx, ry, +1);
        DYN_LAUNCH_CHECK();
    }
    {
        Timed _t(c, "adj_irfft_x", pts);
        launch_irfft_x_out<T>(c, half ? (const T *)pp.spec2 : spec, r2, rx, kc, half);
    }
    DYN_LAUNCH_CHECK();
    {
        Timed _t(c, "adj_div", (double)c->npts);
        hipLaunchKernelGGL(k_adj_div<T>, dim3((unsigned)(((long long)g.NX * g.NYNZ + 255) / 256)), dim3(256), 0, c->stream, g, c->cols,
                           (const T *)r2, uh, vh, wh, step);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

}  // namespace

int pressure_init(Ctx *c) {
    PressurePlan &pp = c->pp;
    const int NZ = c->g.nz;
    pp.ni = c->g.nx - 6;
    pp.nj = c->g.ny;
    pp.nk2 = NZ - 2;
    Radix rx{}, ry{};
    if (!factorize(pp.ni, rx) || !factorize(pp.nj, ry)) {
        set_error("pressure solver: transform sizes need at most 8 radix passes (ni=" + std::to_string(pp.ni) +
                  ", nj=" + std::to_string(pp.nj) + ")");
        return DYN_E_UNSUPPORTED;
    }
    auto smooth = [](const Radix &r) {
        for (int q = 0; q < r.nr; q++)
            if (r.r[q] > 5) return false;
        return true;
    };
    pp.smooth_x = smooth(rx);
    pp.smooth_y = smooth(ry);
    pp.twx = upload_twiddles(pp.ni, c->stream);
    pp.twy = upload_twiddles(pp.nj, c->stream);
    if (!pp.twx || !pp.twy) {
        set_error("pressure solver: twiddle upload failed");
        return DYN_E_HIP;
    }
    DYN_HIP(hipMalloc(&pp.spec, sizeof(double) * 2 * (size_t)pp.ni * pp.nj * pp.nk2));
    int dev = 0, lds = 0;
    DYN_HIP(hipGetDevice(&dev));
    DYN_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    pp.lds_max = (size_t)lds;
    DYN_HIP(hipDeviceGetAttribute(&pp.cus, hipDeviceAttributeMultiprocessorCount, dev));
    // the fused j pass keeps a whole (j, k) slab of one kx in LDS when it fits (FP32 at 256x64, FP64 up to
    // about 192x40) and one thread per real sequence of the solve exists; otherwise the j transforms and
    // the solve run as three kernels
    const size_t need = y_pass_lds(pp.nj, pp.nk2, c->elt);
    const bool f32 = c->prec == DYN_FP32;
    const int nthr = f32 ? ypass_threads<float>() : ypass_threads<double>();
    const int maxb = f32 ? ypass_maxb<float>() : ypass_maxb<double>();
    pp.fused_y = pp.smooth_y && need <= pp.lds_max && 2 * pp.nj <= nthr && max_butterflies(ry, pp.nk2, nthr) <= maxb;
    if (pp.fused_y && f32) {
        DYN_HIP(hipFuncSetAttribute((const void *)k_ypass<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
        DYN_HIP(hipFuncSetAttribute((const void *)k_ypass<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    } else if (pp.fused_y) {
        DYN_HIP(hipFuncSetAttribute((const void *)k_ypass<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
        DYN_HIP(hipFuncSetAttribute((const void *)k_ypass<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    }
    // half slabs (power-of-two nj/2): what a slab of nj/2 rows needs, with one thread per real sequence
    // of the solve
    if (pp.nj % 2 == 0 && pp.nj >= 4 && ((pp.nj / 2) & (pp.nj / 2 - 1)) == 0) {
        Radix rh{};
        const bool ok = pow2_plan48(pp.nj / 2, rh, !f32);
        const size_t needh = y_pass_lds(pp.nj / 2, pp.nk2, c->elt);
        const int hthr = f32 ? half_threads<float>() : half_threads<double>();
        const int hmaxb = f32 ? half_maxb<float>() : half_maxb<double>();
        bool fits = ok && needh <= pp.lds_max && pp.nj <= hthr;
        for (int st = 0; fits && st < rh.nr; st++) {  // butterflies a thread holds in each pass
            const int items = rh.n / rh.r[st] * pp.nk2, per = (items + hthr - 1) / hthr;
            fits = per <= (rh.r[st] == 8 ? pow2_maxb8(hmaxb) : hmaxb);
        }
        pp.half_y = fits && pp.smooth_x;  // its inverse i pass (the row pairs) is an in-place transform
        if (pp.half_y) {
            DYN_HIP(hipMalloc(&pp.spec2, sizeof(double) * 2 * (size_t)pp.ni * pp.nj * pp.nk2));
            if (f32) {
                DYN_HIP(hipFuncSetAttribute((const void *)k_ypass_half<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)needh));
                DYN_HIP(hipFuncSetAttribute((const void *)k_ypass_half<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)needh));
            } else {
                DYN_HIP(hipFuncSetAttribute((const void *)k_ypass_half<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)needh));
                DYN_HIP(hipFuncSetAttribute((const void *)k_ypass_half<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)needh));
            }
        }
    }
    // the power-of-two j transforms of the three-kernel j pass may take more than 64 KiB of LDS
    if (f32) DYN_HIP(hipFuncSetAttribute((const void *)k_fft_y_pow2<float, kFftYMaxB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFftYLds));
    else DYN_HIP(hipFuncSetAttribute((const void *)k_fft_y_pow2<double, kFftYMaxB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFftYLds));
    DYN_HIP(hipMalloc(&pp.ffac, sizeof(double) * (size_t)pp.ni * pp.nj * pp.nk2));
    DYN_HIP(hipMalloc(&pp.zpiv, sizeof(double) * (size_t)pp.ni * pp.nj * pp.nk2));
    DYN_HIP(hipMalloc(&pp.coskx, sizeof(double) * pp.ni));
    DYN_HIP(hipMalloc(&pp.cosly, sizeof(double) * pp.nj));
    DYN_HIP(hipMalloc(&pp.acoef, sizeof(double) * NZ));
    DYN_HIP(hipMalloc(&pp.ccoef, sizeof(double) * NZ));
    // cos(kwave[i]), cos(lwave[j]) exactly as init_fft3d forms them (pressure.cpp:751-752)
    std::vector<double> ck(pp.ni), cl(pp.nj);
    for (int i = 0; i < pp.ni; i++) ck[i] = std::cos((double)i * 2 * kTrigPi / ((double)pp.ni));
    for (int j = 0; j < pp.nj; j++) cl[j] = std::cos((double)j * 2 * kTrigPi / ((double)pp.nj));
    DYN_HIP(hipMemcpy(pp.coskx, ck.data(), sizeof(double) * pp.ni, hipMemcpyHostToDevice));
    DYN_HIP(hipMemcpy(pp.cosly, cl.data(), sizeof(double) * pp.nj, hipMemcpyHostToDevice));
    return DYN_OK;
}

// coefficients that depend on the columns (acoef, ccoef, F2): called once the columns are set
int pressure_coefficients(Ctx *c) {
    PressurePlan &pp = c->pp;
    const int nk = c->g.nz;
    const double dz = c->g.dz;
    const double *mu = c->colhost + (size_t)DYN_COL_MU * nk, *mw = c->colhost + (size_t)DYN_COL_MW * nk;
    const double *rhow = c->colhost + (size_t)DYN_COL_RHOW * nk;
    std::vector<double> a(nk, 0.0), cc(nk, 0.0);
    for (int k = 1; k < nk - 1; k++) {
        a[k] = mu[k] * mw[k] * rhow[k] / (dz * dz);
        cc[k] = mu[k] * mw[k + 1] * rhow[k + 1] / (dz * dz);
    }
    a[1] = 0;
    cc[nk - 2] = 0;
    DYN_HIP(hipMemcpy(pp.acoef, a.data(), sizeof(double) * nk, hipMemcpyHostToDevice));
    DYN_HIP(hipMemcpy(pp.ccoef, cc.data(), sizeof(double) * nk, hipMemcpyHostToDevice));
    int nq = pp.ni * pp.nj;
    hipLaunchKernelGGL(k_thomas_factors, dim3((nq + 127) / 128), dim3(128), 0, c->stream, pp.ni, pp.nj, nk, pp.coskx, pp.cosly,
                       c->cols.rhou, c->cols.rhow, c->cols.mu, c->cols.mw, pp.acoef, pp.ccoef, c->g.dx, dz, pp.ffac, pp.zpiv);
    DYN_LAUNCH_CHECK();
    DYN_HIP(hipStreamSynchronize(c->stream));
    return DYN_OK;
}

int pressure_free(Ctx *c) {
    PressurePlan &pp = c->pp;
    void *ptrs[] = {pp.twx, pp.twy, pp.spec, pp.spec2, pp.ffac, pp.zpiv, pp.coskx, pp.cosly, pp.acoef, pp.ccoef};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    pp = PressurePlan{};
    return DYN_OK;
}

int launch_solve_pressure(Ctx *c, double step, bool halos, PressureHook after_i, void *arg) {
    return c->prec == DYN_FP32 ? solve_t<float>(c, step, halos, after_i, arg) : solve_t<double>(c, step, halos, after_i, arg);
}



// the pressure part of an ensemble member's table entry; false when the member cannot join the batched
// solve (odd ni, or a (j, k) slab too large for the LDS-resident j pass)
bool pressure_ens_fill(Ctx *c, EnsMember &e, size_t &lds_x, size_t &lds_y) {
    PressurePlan &pp = c->pp;
    if (pp.ni % 2 != 0 || !pp.smooth_x || !pp.fused_y || c->prec != DYN_FP32) return false;
    if (!factorize(pp.ni, e.rx) || !factorize(pp.nj, e.ry)) return false;
    e.kcx = x_level_chunk(pp, c->geo, e.rx, sizeof(float), false);
    e.ni = pp.ni;
    e.nj = pp.nj;
    e.nk2 = pp.nk2;
    e.spec = pp.spec;
    e.twx = pp.twx;
    e.twy = pp.twy;
    e.ffac = pp.ffac;
    e.zpiv = pp.zpiv;
    e.coskx = pp.coskx;
    e.cosly = pp.cosly;
    e.acoef = pp.acoef;
    e.dx = c->g.dx;
    e.dz = c->g.dz;
    lds_x = x_pass_lds(pp.ni, e.kcx, sizeof(float));
    lds_y = y_pass_lds(pp.nj, pp.nk2, sizeof(float));
    return lds_y <= pp.lds_max && 2 * pp.nj <= 1024 && max_butterflies(e.ry, pp.nk2, 1024) <= kMaxBY;
}

// solve_pressure(step) of every member without the uvw halo copy (as in the fused step): three launches
int launch_pressure_ens(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, double step, int tot_px, int tot_py, size_t lds_x,
                        size_t lds_y) {
    static size_t lds_y_set = 0;
    if (lds_y > lds_y_set) {
        DYN_HIP(hipFuncSetAttribute((const void *)k_ens_ypass<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_y));
        lds_y_set = lds_y;
    }
    {
        const int grid = tot_px;
        Timed _t(c0, "ens_p1_div_rfft_x", 0.0);
        hipLaunchKernelGGL(k_ens_div_rfft_x<float>, dim3(grid), dim3(kThreadsX), lds_x, c0->stream, d, nmem, ro, tot_px, step);
        DYN_LAUNCH_CHECK();
    }
    {
        const int grid = tot_py;
        Timed _t(c0, "ens_p2_ypass", 0.0);
        hipLaunchKernelGGL(k_ens_ypass<float>, dim3(grid), dim3(1024), lds_y, c0->stream, d, nmem, tot_py);
        DYN_LAUNCH_CHECK();
    }
    {
        const int grid = tot_px;
        Timed _t(c0, "ens_p5_irfft_x_out", 0.0);
        hipLaunchKernelGGL(k_ens_irfft_x_out<float>, dim3(grid), dim3(kThreadsX), lds_x, c0->stream, d, nmem, tot_px);
        DYN_LAUNCH_CHECK();
    }
    return DYN_OK;
}

// adjoint of solve_pressure(step) minus its periodic_uvw_boundaries: consumes (and zeroes) the adjoint of
// PI and adds the adjoint of the divergence input to the forecast-velocity adjoints; r1, r2 are scratch
// fields of the context's element type
int launch_pressure_adjoint(Ctx *c, double step, void *pihat, void *uh, void *vh, void *wh, void *r1, void *r2) {
    if (c->prec == DYN_FP32)
        return solve_adjoint_t<float>(c, step, (float *)pihat, (float *)uh, (float *)vh, (float *)wh, (float *)r1, (float *)r2);
    return solve_adjoint_t<double>(c, step, (double *)pihat, (double *)uh, (double *)vh, (double *)wh, (double *)r1, (double *)r2);
}

}  // namespace dyn
