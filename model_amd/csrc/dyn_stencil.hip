// Advection, pressure-gradient, boundary, Kessler and Rayleigh kernels for gfx950.
//
// Each kernel restates one reference operator over exactly the index ranges the reference writes.
// Fields are [i][j][k] with k fastest: a 64-lane wavefront spans 64 consecutive levels of one (i,j)
// column, so every neighbour load in i or j is a coalesced 64-wide row and every k-neighbour load
// reuses the lines the wavefront already pulled into L1/L2.
#include "dyn_internal.hpp"

namespace dyn {
namespace {

#define IX(i, j, k) ((long long)(i) * g.NYNZ + (long long)(j) * g.NZ + (long long)(k))

// ---- face interpolants, fluxes.cpp:110-206 ---------------------------------------------------
template <typename T>
__device__ __forceinline__ T interp5(const T *p, long long st, T sign) {
    return face_i5(p[-2 * st], p[-st], p[0], p[st], p[2 * st], p[3 * st], sign);
}
template <typename T>
__device__ __forceinline__ T interp3(const T *p, T sign) { return face_i3(p[-1], p[0], p[1], p[2], sign); }
template <typename T>
__device__ __forceinline__ T interp2(const T *p) { return T(0.5) * (p[1] + p[0]); }
// signof, fluxes.cpp:295-311 (zero maps to -1)
template <typename T>
__device__ __forceinline__ T signof(T x) { return x > T(0) ? T(1) : T(-1); }

struct VelFields {
    const void *U, *V, *W, *UB, *VB, *WB;
};

// ---- momentum control-volume faces (fluxes.cpp:1109-1592, compute_fluxes 582-627) -------------
// u control volume: east (i+1/2), north (j+1/2), top (k+1/2)
template <typename T>
__device__ __forceinline__ T u_east(const Geo &g, const T *U, const T *UB, int i, int j, int k) {
    T ua = T(0.5) * (U[IX(i + 1, j, k)] + U[IX(i, j, k)]);
    T uba = T(0.5) * (UB[IX(i + 1, j, k)] + UB[IX(i, j, k)]);
    T ue = interp5(U + IX(i, j, k), g.NYNZ, signof(uba + ua));
    T ube = interp5(UB + IX(i, j, k), g.NYNZ, signof(ua));
    return ua * ube + (uba + ua) * ue;
}
template <typename T>
__device__ __forceinline__ T u_north(const Geo &g, const T *U, const T *V, const T *UB, const T *VB, int i, int j, int k) {
    T va = T(0.5) * (V[IX(i, j + 1, k)] + V[IX(i - 1, j + 1, k)]);
    T vba = T(0.5) * (VB[IX(i, j + 1, k)] + VB[IX(i - 1, j + 1, k)]);
    T un = interp5(U + IX(i, j, k), (long long)g.NZ, signof(vba + va));
    T ubn = interp5(UB + IX(i, j, k), (long long)g.NZ, signof(va));
    return va * ubn + (vba + va) * un;
}
template <typename T>
__device__ __forceinline__ T u_top(const Geo &g, const T *U, const T *W, const T *UB, const T *WB, int i, int j, int k) {
    if (k == 0) return T(0);
    T wa = T(0.5) * (W[IX(i, j, k + 1)] + W[IX(i - 1, j, k + 1)]);
    T wba = T(0.5) * (WB[IX(i, j, k + 1)] + WB[IX(i - 1, j, k + 1)]);
    T ut, ubt;
    if (k == 1 || k == g.NZ - 2) {
        ut = interp2(U + IX(i, j, k));
        ubt = interp2(UB + IX(i, j, k));
    } else {
        ut = interp3(U + IX(i, j, k), signof(wba + wa));
        ubt = interp3(UB + IX(i, j, k), signof(wa));
    }
    return wa * ubt + (wba + wa) * ut;
}
// v control volume
template <typename T>
__device__ __forceinline__ T v_east(const Geo &g, const T *U, const T *V, const T *UB, const T *VB, int i, int j, int k) {
    T ua = T(0.5) * (U[IX(i + 1, j, k)] + U[IX(i + 1, j - 1, k)]);
    T uba = T(0.5) * (UB[IX(i + 1, j, k)] + UB[IX(i + 1, j - 1, k)]);
    T ve = interp5(V + IX(i, j, k), g.NYNZ, signof(uba + ua));
    T vbe = interp5(VB + IX(i, j, k), g.NYNZ, signof(ua));
    return ua * vbe + (uba + ua) * ve;
}
template <typename T>
__device__ __forceinline__ T v_north(const Geo &g, const T *V, const T *VB, int i, int j, int k) {
    T va = T(0.5) * (V[IX(i, j + 1, k)] + V[IX(i, j, k)]);
    T vba = T(0.5) * (VB[IX(i, j + 1, k)] + VB[IX(i, j, k)]);
    T vn = interp5(V + IX(i, j, k), (long long)g.NZ, signof(vba + va));
    T vbn = interp5(VB + IX(i, j, k), (long long)g.NZ, signof(va));
    return va * vbn + (vba + va) * vn;
}
template <typename T>
__device__ __forceinline__ T v_top(const Geo &g, const T *V, const T *W, const T *VB, const T *WB, int i, int j, int k) {
    if (k == 0) return T(0);
    T wa = T(0.5) * (W[IX(i, j, k + 1)] + W[IX(i, j - 1, k + 1)]);
    T wba = T(0.5) * (WB[IX(i, j, k + 1)] + WB[IX(i, j - 1, k + 1)]);
    T vt, vbt;
    if (k == 1 || k == g.NZ - 2) {
        vt = interp2(V + IX(i, j, k));
        vbt = interp2(VB + IX(i, j, k));
    } else {
        vt = interp3(V + IX(i, j, k), signof(wba + wa));
        vbt = interp3(VB + IX(i, j, k), signof(wa));
    }
    return wa * vbt + (wba + wa) * vt;
}

// ---- scalar faces with a basic state (interpolate_scalar 797-942 + compute_fluxes_scalar 632-654)
template <typename T>
__device__ __forceinline__ T s_east(const Geo &g, const T *S, const T *SB, const T *U, const T *UB, int i, int j, int k) {
    T U1 = U[IX(i + 1, j, k)], UB1 = UB[IX(i + 1, j, k)];
    T se = interp5(S + IX(i, j, k), g.NYNZ, signof(UB1 + U1));
    T be = interp5(SB + IX(i, j, k), g.NYNZ, signof(U1));
    return U1 * be + (UB1 + U1) * se;
}
template <typename T>
__device__ __forceinline__ T s_north(const Geo &g, const T *S, const T *SB, const T *V, const T *VB, int i, int j, int k) {
    T V1 = V[IX(i, j + 1, k)], VB1 = VB[IX(i, j + 1, k)];
    T sn = interp5(S + IX(i, j, k), (long long)g.NZ, signof(VB1 + V1));
    T bn = interp5(SB + IX(i, j, k), (long long)g.NZ, signof(V1));
    return V1 * bn + (VB1 + V1) * sn;
}
template <typename T>
__device__ __forceinline__ T s_top(const Geo &g, const T *S, const T *SB, const T *W, const T *WB, int i, int j, int k) {
    if (k == 0) return T(0);
    T W1 = W[IX(i, j, k + 1)], WB1 = WB[IX(i, j, k + 1)];
    T st, bt;
    if (k == 1 || k == g.NZ - 2) {
        st = interp2(S + IX(i, j, k));
        bt = interp2(SB + IX(i, j, k));
    } else {
        st = interp3(S + IX(i, j, k), signof(WB1 + W1));
        bt = interp3(SB + IX(i, j, k), signof(W1));
    }
    return W1 * bt + (WB1 + W1) * st;
}
// perturbation-only scalar faces (qc, qr)
template <typename T>
__device__ __forceinline__ T q_east(const Geo &g, const T *S, const T *U, const T *UB, int i, int j, int k) {
    T ub = UB[IX(i + 1, j, k)] + U[IX(i + 1, j, k)];
    return ub * interp5(S + IX(i, j, k), g.NYNZ, signof(ub));
}
template <typename T>
__device__ __forceinline__ T q_north(const Geo &g, const T *S, const T *V, const T *VB, int i, int j, int k) {
    T vb = VB[IX(i, j + 1, k)] + V[IX(i, j + 1, k)];
    return vb * interp5(S + IX(i, j, k), (long long)g.NZ, signof(vb));
}
template <typename T>
__device__ __forceinline__ T qc_top(const Geo &g, const T *S, const T *W, const T *WB, int i, int j, int k) {
    if (k == 0) return T(0);
    T wb = WB[IX(i, j, k + 1)] + W[IX(i, j, k + 1)];
    T t = (k == 1 || k == g.NZ - 2) ? interp2(S + IX(i, j, k)) : interp3(S + IX(i, j, k), signof(wb));
    return wb * t;
}

// rain fall speed at level k of one column (kessler.cpp:222-252); level 0 copies level 1,
// level NZ-1 is never written by the reference and is read from the array
template <typename T>
__device__ __forceinline__ T fall_speed(const Geo &g, const Cols &c, const T *QR, const T *VT, long long col, int k) {
    if (k == g.NZ - 1) return VT[col + k];
    int kk = k == 0 ? 1 : k;
    T qr = QR[col + kk];
    T vt = T(0);
    if (qr > T(1.0e-12)) {
        T rk = T(c.rhou[kk]);
        T qrr = fmax(qr * T(0.001) * rk, T(0));
        T vtden = sqrt(T(c.rhou[1]) / rk);
        vt = T(36.34) * (pow(qrr, T(0.1364))) * vtden;
    }
    return vt;
}
template <typename T>
__device__ __forceinline__ T qr_top(const Geo &g, const Cols &c, const T *QR, const T *VT, const T *W, const T *WB, int i, int j, int k) {
    long long col = IX(i, j, 0);
    if (k == 0) {
        T t0 = T(0.5) * (QR[col] + QR[col + 1]);
        return -T(0.5) * (fall_speed(g, c, QR, VT, col, 1) + fall_speed(g, c, QR, VT, col, 0)) * t0;
    }
    T vk = fall_speed(g, c, QR, VT, col, k), vk1 = fall_speed(g, c, QR, VT, col, k + 1);
    T wb = WB[col + k + 1] + W[col + k + 1];
    T t;
    if (k == 1 || k == g.NZ - 2) t = interp2(QR + col + k);
    else t = interp3(QR + col + k, signof(WB[col + k + 1] + W[col + k + 1] - T(0.5) * (vk + vk1)));
    return (wb - T(0.5) * (vk1 + vk)) * t;
}

// ---- advect_uvw_velocity (advection.cpp:304-339) ------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_advect_uvw(Geo g, Cols c, const T *__restrict__ U, const T *__restrict__ V,
                                                    const T *__restrict__ W, const T *__restrict__ UB, const T *__restrict__ VB,
                                                    const T *__restrict__ WB, const T *__restrict__ UM, const T *__restrict__ VM,
                                                    const T *__restrict__ WM, const T *__restrict__ TH, const T *__restrict__ QV,
                                                    const T *__restrict__ QC, const T *__restrict__ QR, T *__restrict__ UP,
                                                    T *__restrict__ VP, T *__restrict__ WP, T step, int il, int jl, int jh, int moist) {
    int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + jl;
    int i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    const T dtx = T(g.dtx), dty = T(g.dty), dt = T(g.dt);
    const T DTZ = T(g.dtz) * T(c.mu[k]);
    const T rw1 = T(c.rhow[k + 1]), rw0 = T(c.rhow[k]), odr = T(c.one_d_rhou[k]), fj = T(c.f[j]);
    // u_tend_cell, advection.cpp:73-94
    {
        T e = u_east(g, U, UB, i, j, k), w = u_east(g, U, UB, i - 1, j, k);
        T n = u_north(g, U, V, UB, VB, i, j, k), s = u_north(g, U, V, UB, VB, i, j - 1, k);
        T t = u_top(g, U, W, UB, WB, i, j, k), b = u_top(g, U, W, UB, WB, i, j, k - 1);
        T other = T(0.25) * (V[IX(i, j, k)] + V[IX(i, j + 1, k)] + V[IX(i - 1, j, k)] + V[IX(i - 1, j + 1, k)]);
        T tend = -dtx * (e - w) - dty * (n - s) - DTZ * (rw1 * t - rw0 * b) * odr + dt * fj * other;
        UP[IX(i, j, k)] = UM[IX(i, j, k)] + step * tend;
    }
    // v_tend_cell, advection.cpp:101-122
    {
        T e = v_east(g, U, V, UB, VB, i, j, k), w = v_east(g, U, V, UB, VB, i - 1, j, k);
        T n = v_north(g, V, VB, i, j, k), s = v_north(g, V, VB, i, j - 1, k);
        T t = v_top(g, V, W, VB, WB, i, j, k), b = v_top(g, V, W, VB, WB, i, j, k - 1);
        T other = T(0.25) * (U[IX(i, j, k)] + U[IX(i + 1, j, k)] + U[IX(i, j - 1, k)] + U[IX(i + 1, j - 1, k)]);
        T tend = -dtx * (e - w) - dty * (n - s) - DTZ * (rw1 * t - rw0 * b) * odr - dt * fj * other;
        VP[IX(i, j, k)] = VM[IX(i, j, k)] + step * tend;
    }
    // buoyancy overwrite of w, advection.cpp:329-334
    if (k >= 2) {
        T bk, bkm;
        if (!moist) {
            bk = TH[IX(i, j, k)] / T(c.tbv[k]);
            bkm = TH[IX(i, j, k - 1)] / T(c.tbv[k - 1]);
        } else {
            bk = TH[IX(i, j, k)] / T(c.tbv[k]) + T(0.61) * QV[IX(i, j, k)] - QC[IX(i, j, k)] - QR[IX(i, j, k)];
            bkm = TH[IX(i, j, k - 1)] / T(c.tbv[k - 1]) + T(0.61) * QV[IX(i, j, k - 1)] - QC[IX(i, j, k - 1)] - QR[IX(i, j, k - 1)];
        }
        WP[IX(i, j, k)] = WM[IX(i, j, k)] + step * dt * T(kGrav) * T(0.5) * (bk + bkm);
    }
}

// Z_ADVECT_BASE, advection.cpp:16-18
template <typename T>
__device__ __forceinline__ T z_advect_base(const Geo &g, const Cols &c, const double *base, const T *W, int i, int j, int k) {
    return (T(0.50) * (T(c.rhow[k + 1]) * W[IX(i, j, k + 1)] * (T(base[k + 1]) - T(base[k])) +
                       T(c.rhow[k]) * W[IX(i, j, k)] * (T(base[k]) - T(base[k - 1]))) *
            T(c.one_d_rhou[k]));
}

// ---- advect_theta (advection.cpp:349-373) -----------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_advect_theta(Geo g, Cols c, const T *__restrict__ U, const T *__restrict__ V,
                                                      const T *__restrict__ W, const T *__restrict__ UB, const T *__restrict__ VB,
                                                      const T *__restrict__ WB, const T *__restrict__ TH, const T *__restrict__ THB,
                                                      const T *__restrict__ THM, T *__restrict__ THP, T step, int il, int jl, int jh) {
    int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + jl;
    int i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    const T DTZ = T(g.dtz) * T(c.mu[k]);
    T e = s_east(g, TH, THB, U, UB, i, j, k), w = s_east(g, TH, THB, U, UB, i - 1, j, k);
    T n = s_north(g, TH, THB, V, VB, i, j, k), s = s_north(g, TH, THB, V, VB, i, j - 1, k);
    T t = s_top(g, TH, THB, W, WB, i, j, k), b = s_top(g, TH, THB, W, WB, i, j, k - 1);
    T tend = -T(g.dtx) * (e - w) - T(g.dty) * (n - s) - DTZ * (T(c.rhow[k + 1]) * t - T(c.rhow[k]) * b) * T(c.one_d_rhou[k]) -
             DTZ * z_advect_base(g, c, c.tb, W, i, j, k);
    THP[IX(i, j, k)] = THM[IX(i, j, k)] + step * tend;
}

// ---- advect_microphysics_cell (advection.cpp:451-498) -------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_advect_moisture(Geo g, Cols c, const T *__restrict__ U, const T *__restrict__ V,
                                                         const T *__restrict__ W, const T *__restrict__ UB, const T *__restrict__ VB,
                                                         const T *__restrict__ WB, const T *__restrict__ QV, const T *__restrict__ QC,
                                                         const T *__restrict__ QR, const T *__restrict__ QB, const T *__restrict__ QVM,
                                                         const T *__restrict__ QCM, const T *__restrict__ QRM, T *__restrict__ QVP,
                                                         T *__restrict__ QCP, T *__restrict__ QRP, T *__restrict__ VT, T step, int il,
                                                         int jl, int jh) {
    int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + jl;
    int i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    const T DTZ = T(g.dtz) * T(c.mu[k]);
    const T dtx = T(g.dtx), dty = T(g.dty), rw1 = T(c.rhow[k + 1]), rw0 = T(c.rhow[k]), odr = T(c.one_d_rhou[k]);
    // vapour, qv_tend_cell advection.cpp:197-213
    {
        T e = s_east(g, QV, QB, U, UB, i, j, k), w = s_east(g, QV, QB, U, UB, i - 1, j, k);
        T n = s_north(g, QV, QB, V, VB, i, j, k), s = s_north(g, QV, QB, V, VB, i, j - 1, k);
        T t = s_top(g, QV, QB, W, WB, i, j, k), b = s_top(g, QV, QB, W, WB, i, j, k - 1);
        T tend = -dtx * (e - w) - dty * (n - s) - DTZ * (rw1 * t - rw0 * b) * odr - DTZ * z_advect_base(g, c, c.qb, W, i, j, k);
        QVP[IX(i, j, k)] = QVM[IX(i, j, k)] + step * tend;
    }
    // cloud, qc_tend_cell advection.cpp:220-233
    {
        T e = q_east(g, QC, U, UB, i, j, k), w = q_east(g, QC, U, UB, i - 1, j, k);
        T n = q_north(g, QC, V, VB, i, j, k), s = q_north(g, QC, V, VB, i, j - 1, k);
        T t = qc_top(g, QC, W, WB, i, j, k), b = qc_top(g, QC, W, WB, i, j, k - 1);
        T tend = -dtx * (e - w) - dty * (n - s) - DTZ * (rw1 * t - rw0 * b) * odr;
        QCP[IX(i, j, k)] = QCM[IX(i, j, k)] + step * tend;
    }
    // rain, qr_tend_cell advection.cpp:240-253 (fall speed computed on the fly from QR)
    {
        T e = q_east(g, QR, U, UB, i, j, k), w = q_east(g, QR, U, UB, i - 1, j, k);
        T n = q_north(g, QR, V, VB, i, j, k), s = q_north(g, QR, V, VB, i, j - 1, k);
        T t = qr_top(g, c, QR, (const T *)VT, W, WB, i, j, k), b = qr_top(g, c, QR, (const T *)VT, W, WB, i, j, k - 1);
        T tend = -dtx * (e - w) - dty * (n - s) - DTZ * (rw1 * t - rw0 * b) * odr;
        QRP[IX(i, j, k)] = QRM[IX(i, j, k)] + step * tend;
    }
}

// the stored fall-speed field (calculate_rain_fall_speed_kessler writes it; precip_rate reads it)
template <typename T>
__global__ void k_fall_speed(Geo g, Cols c, const T *__restrict__ QR, T *__restrict__ VT, int il, int jl, int jh) {
    int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + jl;
    int i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    long long col = IX(i, j, 0);
    T v = fall_speed(g, c, QR, (const T *)VT, col, k);
    VT[col + k] = v;
    if (k == 1) VT[col] = v;
}

// ---- zero_moisture (microphysics.cpp:675-700) ---------------------------------------------------
template <typename T>
__global__ void k_zero_moisture(Geo g, Cols c, T *__restrict__ QVP, T *__restrict__ QCP, T *__restrict__ QRP,
                                const T *__restrict__ QB, long long n, int il, int ih, int jl, int jh) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (QCP[p] < T(0)) QCP[p] = T(0);
    if (QRP[p] < T(0)) QRP[p] = T(0);
    int k = (int)(p % g.NZ);
    long long r = p / g.NZ;
    int j = (int)(r % g.NY), i = (int)(r / g.NY);
    if (i >= il && i < ih && j >= jl && j < jh && k >= 1 && k < g.NZ - 1) {
        if (QVP[p] + QB[p] + T(c.qb[k]) < T(0)) QVP[p] = -(QB[p] + T(c.qb[k]));
    }
}

// ---- pressure-gradient force, solver.cpp:123-135 ------------------------------------------------
template <typename T>
__global__ void k_pgf(Geo g, Cols c, const T *__restrict__ PI, T *__restrict__ UP, T *__restrict__ VP, T *__restrict__ WP, T step,
                      int il, int jl, int jh) {
    int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + jl;
    int i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    long long p = IX(i, j, k);
    UP[p] -= step * (T(g.dtx) * (PI[p] - PI[p - g.NYNZ]));
    VP[p] -= step * (T(g.dty) * (PI[p] - PI[p - g.NZ]));
    if (k >= 2) WP[p] -= step * ((T(g.dtz) * T(c.mw[k])) * (PI[p] - PI[p - 1]));
}

// ---- boundaries (serial periodic branch), boundaries.cpp:162-191, 254-264 ------------------------
struct FieldList {
    void *f[5];
    int n;
};

// sponge_boundaries_north_south: rows [0,js) and [NY-jn,NY), i in [1,NX), k in [1,NZ)
template <typename T>
__global__ void k_sponge_ns(Geo g, FieldList fl, const double *__restrict__ cs, const double *__restrict__ cn, int js, int jn) {
    int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int r = blockIdx.y;  // band row: 0..js-1 south, js..js+jn-1 north
    int i = blockIdx.z + 1;
    if (k >= g.NZ) return;
    int j;
    T coef;
    if (r < js) { j = r; coef = T(cs[r]); }
    else { j = g.NY - jn + (r - js); coef = T(cn[r - js]); }
    long long p = IX(i, j, k);
    for (int f = 0; f < fl.n; f++) {
        T *A = (T *)fl.f[f];
        A[p] -= coef * A[p];
    }
}

// periodic_boundaries_east_west: six halo columns copied for every (j,k)
template <typename T>
__global__ void k_periodic(Geo g, FieldList fl) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.NYNZ) return;
    const long long s = g.NYNZ;
    for (int f = 0; f < fl.n; f++) {
        T *A = (T *)fl.f[f];
        A[q + 0 * s] = A[q + (long long)(g.NX - 6) * s];
        A[q + 1 * s] = A[q + (long long)(g.NX - 5) * s];
        A[q + 2 * s] = A[q + (long long)(g.NX - 4) * s];
        A[q + (long long)(g.NX - 3) * s] = A[q + 3 * s];
        A[q + (long long)(g.NX - 2) * s] = A[q + 4 * s];
        A[q + (long long)(g.NX - 1) * s] = A[q + 5 * s];
    }
}

// upper_lower_boundaries(var,1,NX,1,NY) for nfill fields, then upper_lower_boundaries_zero for nzero
template <typename T>
__global__ void k_upper_lower(Geo g, FieldList fl, int nfill) {
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int i = blockIdx.y + 1;
    if (j >= g.NY) return;
    long long p = IX(i, j, 0);
    for (int f = 0; f < fl.n; f++) {
        T *A = (T *)fl.f[f];
        if (f < nfill) {
            A[p] = A[p + 1];
            A[p + g.NZ - 1] = A[p + g.NZ - 2];
        } else {
            A[p] = T(0);
            A[p + 1] = T(0);
            A[p + g.NZ - 1] = T(0);
        }
    }
}

// ---- Kessler warm rain, kessler.cpp:25-183 -------------------------------------------------------
// The step-end boundary update of theta and moisture in the columns it changes only at the lid, done
// by Kessler's threads of levels 1 and NZ-2 (the lid takes their values), with theta's Rayleigh damping
// from level `lev` up (after the lid, as the reference orders them): columns i in [i0, i1), j in [j0, j1).
// qr's bottom lid waits for k_precip, which reads level 0 before it writes the lid.
struct KessTail {
    int on = 0;
    int i0 = 0, i1 = 0, j0 = 0, j1 = 0;
    int lev = 1 << 30;
    const double *ray = nullptr;
    __device__ bool in(int i, int j) const { return on && i >= i0 && i < i1 && j >= j0 && j < j1; }
};
// x / y in Kessler's FP64 thermodynamics: the division itself in the FP64 form (the reference's
// arithmetic); in the FP32 form, whose results are rounded to FP32, x times the hardware reciprocal refined
// by two Newton steps (within an ulp of the FP64 quotient, five dependent operations where the correctly
// rounded division takes ten; the kernel is bound by instruction issue)
template <typename T>
__device__ __forceinline__ double kdiv(double x, double y) {
    if constexpr (sizeof(T) == 8) {
        return x / y;
    } else {
        double r = __builtin_amdgcn_rcp(y);
        double e = fma(-y, r, 1.0);
        r = fma(r, e, r);
        e = fma(-y, r, 1.0);
        r = fma(r, e, r);
        return x * r;
    }
}

// pressure^(cp/Rd) of the Exner function pressure: the reference's pow in the FP64 form.  In the FP32 form
// pressure^3 * sqrt(pressure) * exp(t), t = (cp/Rd - 3.5) ln(pressure) = -0.00174 ln(pressure): ln from the
// binary exponent and 2 atanh(s) of the mantissa m in [1/sqrt2, sqrt2), s = (m-1)/(m+1), |s| < 0.172,
// series to s^15 (remainder 1e-14); exp(t) to t^5 (|t| < 5e-3 for pressure > 0.05).  Within 4 ulps of the
// exact power over pressure in [0.05, 1.3] (pow: 0.7), so the FP32 results match pow's but for rare
// roundings; about 60 instructions where pow takes 250 (the kernel is bound by instruction issue)
template <typename T>
__device__ __forceinline__ double kpow_cpRd(double x) {
    if constexpr (sizeof(T) == 8) {
        return pow(x, kCp / kRd);
    } else {
        int ex;
        double m = frexp(x, &ex);
        if (m < 0.70710678118654752) {
            m *= 2.0;
            ex -= 1;
        }
        const double s = kdiv<T>(m - 1.0, m + 1.0), s2 = s * s;
        double q = 1.0 / 15;
        q = fma(q, s2, 1.0 / 13);
        q = fma(q, s2, 1.0 / 11);
        q = fma(q, s2, 1.0 / 9);
        q = fma(q, s2, 1.0 / 7);
        q = fma(q, s2, 1.0 / 5);
        q = fma(q, s2, 1.0 / 3);
        q = fma(q, s2, 1.0);
        const double lnx = fma(double(ex), 0.6931471805599453, 2.0 * s * q);
        const double t = (kCp / kRd - 3.5) * lnx;
        double e = fma(t, 1.0 / 120, 1.0 / 24);
        e = fma(e, t, 1.0 / 6);
        e = fma(e, t, 0.5);
        e = fma(e, t, 1.0);
        e = fma(e, t, 1.0);
        return (x * x) * (x * sqrt(x)) * e;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_kessler(Geo g, Cols c, T *__restrict__ THP, T *__restrict__ QVP, T *__restrict__ QCP,
                                                 T *__restrict__ QRP, const T *__restrict__ THB, const T *__restrict__ QB,
                                                 const T *__restrict__ PB, const T *__restrict__ PI, int *__restrict__ flag,
                                                 unsigned char *__restrict__ sat, int il, int jl, int jh, KessTail tail) {
    int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + jl;
    int i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    long long p = IX(i, j, k);
    // Thermodynamics (full theta, pressure, vapour, temperature, saturation and the condensate) in FP64
    // registers in both forms, rates in the storage type.  The condensate and the evaporation are small
    // differences of ~1e-2 mixing ratios whose saturation value moves 6 % per kelvin, so in FP32 the
    // rounding of the 300 K temperature (3e-5 K) alone would set their error.  In the FP64 form R = T and
    // every expression is the reference's, in its order.
    using R = double;
    const R cp = R(kCp), Lv = R(kLv), tbk = R(c.tb[k]), qbk = R(c.qb[k]);
    const T dt = T(g.dt), rhou = T(c.rhou[k]);
    T th = THP[p], qv = QVP[p], qc = QCP[p], qr = QRP[p];
    R theta = R(th) + R(THB[p]) + tbk;
    R pressure = kdiv<T>(R(PI[p]), cp * R(c.tbv[k])) + R(PB[p]);
    R vapor = R(qv) + R(QB[p]) + qbk;
    R temperature = theta * pressure;
    R pd = R(kP0) * kpow_cpRd<T>(pressure);
    T A = qc > T(1.0e-3) ? T(0.001) * (qc - T(1.0e-3)) * dt : T(0);
    T B = qr > T(1.0e-12) ? rhou * T(2.2) * qc * pow(qr, T(0.875)) * dt : T(0);
    T E = T(0);
    if (qr > T(1.0e-12)) {
        R esl = R(611.2) * exp(kdiv<T>(R(17.67) * (temperature - R(273.15)), temperature - R(29.65)));
        R qvsat = kdiv<T>(R(0.62197) * esl, pd - esl);
        T cvent = T(1.6) + T(30.3922) * pow(rhou * qr, T(0.2046));
        R e = kdiv<T>(R(1.0), R(rhou)) * kdiv<T>((R(1.0) - kdiv<T>(vapor, qvsat)) * R(cvent) * R(pow(rhou * qr, T(0.525))),
                                                R(2.03e4) + kdiv<T>(R(9.58e6), pd * qvsat)) * R(dt);
        E = T(e);
        E = fmin(qr, E);
        E = fmax(E, T(0));
        if (E > T(100)) atomicOr(flag, 1);
    }
    qv += E;
    T total_convert = fmin(A + B, qc);
    qc -= total_convert;
    qr += (total_convert - E);
    R de = kdiv<T>(Lv, cp * pressure) * R(E);
    th -= T(de);
    theta = R(th) + R(THB[p]) + tbk;
    vapor = R(qv) + R(QB[p]) + qbk;
    temperature = theta * pressure;
    R esl = R(611.2) * exp(kdiv<T>(R(17.67) * (temperature - R(273.15)), temperature - R(29.65)));
    R qvsat = kdiv<T>(R(0.62197) * esl, pd - esl);
    const R f = R(17.67 * (273.15 - 29.65) * 2.5e6 / 1004.0);
    R tm = temperature - R(29.65);
    R phi = kdiv<T>(kdiv<T>(pd, pd - esl) * qvsat * f, tm * tm);
    R C = R(0);
    // sat: isSaturated of kessler.cpp:129-150, for the turbulence's moist stability (nullptr when off)
    if (vapor > qvsat) {
        C = kdiv<T>(vapor - qvsat, R(1.0) + phi);
        if (sat) sat[p] = 1;
    } else if (vapor < qvsat && qc > T(0)) {
        C = -fmin(R(qc), kdiv<T>(qvsat - vapor, R(1.0) + phi));
        if (sat) sat[p] = -C < R(qc) ? 1 : 0;
    } else if (sat) {
        sat[p] = vapor == qvsat ? 1 : 0;
    }
    qc += T(C);
    qv -= T(C);
    R dc = kdiv<T>(Lv, cp * pressure) * C;
    th += T(dc);
    if (tail.in(i, j)) {
        auto damp = [&](int kk, T v) { return kk >= tail.lev ? v - T(tail.ray[kk]) * v : v; };
        const int kl = k == 1 ? 0 : (k == g.NZ - 2 ? g.NZ - 1 : -1);  // the lid level this level feeds
        if (kl >= 0) {
            const long long pl = p + (kl - k);
            THP[pl] = damp(kl, th);
            QVP[pl] = qv;
            QCP[pl] = qc;
            if (kl != 0) QRP[pl] = qr;
        }
        th = damp(k, th);
    }
    THP[p] = th;
    QVP[p] = qv;
    QCP[p] = qc;
    QRP[p] = qr;
}

// precip_rate, microphysics.cpp:143-160
template <typename T>
__global__ void k_precip(Geo g, Cols c, const T *__restrict__ VT, T *__restrict__ QRP, T *__restrict__ acc, int il, int ih,
                         int jl, int jh, KessTail tail) {
    int j = blockIdx.x * blockDim.x + threadIdx.x + jl;
    int i = blockIdx.y + il;
    if (j >= jh || i >= ih) return;
    long long p = IX(i, j, 0);
    acc[(long long)i * g.NY + j] += T(c.rhow[1]) * T(0.5) * (VT[p] + VT[p + 1]) * T(0.5) * (QRP[p] + QRP[p + 1]) * T(g.dt);
    if (tail.in(i, j)) QRP[p] = QRP[p + 1];  // the step-end lid of qr, after the rate read level 0
}

// rayleigh_damping, damping.cpp:1148-1177; coefficients [0..NZ) for u/v/theta, [NZ..2NZ) for w
template <typename T>
__global__ void k_rayleigh(Geo g, const double *__restrict__ coef, T *__restrict__ UP, T *__restrict__ VP, T *__restrict__ WP,
                           T *__restrict__ THP, int il, int jl, int jh, int level) {
    int k = blockIdx.x * blockDim.x + threadIdx.x + level;
    int j = blockIdx.y * blockDim.y + threadIdx.y + jl;
    int i = blockIdx.z + il;
    if (k >= g.NZ || j >= jh) return;
    long long p = IX(i, j, k);
    T cu = T(coef[k]), cw = T(coef[g.NZ + k]);
    UP[p] -= cu * UP[p];
    VP[p] -= cu * VP[p];
    WP[p] -= cw * WP[p];
    THP[p] -= cu * THP[p];
}

template <typename T>
__global__ void k_convert(T *dst, const double *src, long long n) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = T(src[p]);
}
template <typename T>
__global__ void k_convert_back(double *dst, const T *src, long long n) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = double(src[p]);
}

inline dim3 col_grid(const Geo &g, int nk, int jl, int jh, int il, int ih, dim3 &blk) {
    blk = dim3(64, 4, 1);
    return dim3((nk + 63) / 64, (jh - jl + 3) / 4, ih - il);
}

template <typename T>
int advect_uvw_t(Ctx *c, double step, int il, int ih, int jl, int jh) {
    dim3 blk, grd = col_grid(c->geo, c->geo.NZ - 2, jl, jh, il, ih, blk);
    bool moist = c->g.microphysics != 0;
    auto P = [&](int f, int r) { return (T *)c->prog[f][r]; };
    {
        Timed _t(c, "advect_uvw", (double)(ih - il) * (jh - jl) * (c->geo.NZ - 2));
        hipLaunchKernelGGL(k_advect_uvw<T>, grd, blk, 0, c->stream, c->geo, c->cols, P(DYN_U, 1), P(DYN_V, 1), P(DYN_W, 1),
                       (const T *)c->bar[0], (const T *)c->bar[1], (const T *)c->bar[2], P(DYN_U, 0), P(DYN_V, 0), P(DYN_W, 0),
                       P(DYN_TH, 1), moist ? P(DYN_QV, 1) : nullptr, moist ? P(DYN_QC, 1) : nullptr, moist ? P(DYN_QR, 1) : nullptr,
                       P(DYN_U, 2), P(DYN_V, 2), P(DYN_W, 2), T(step), il, jl, jh, moist ? 1 : 0);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int advect_theta_t(Ctx *c, double step, int il, int ih, int jl, int jh) {
    dim3 blk, grd = col_grid(c->geo, c->geo.NZ - 2, jl, jh, il, ih, blk);
    auto P = [&](int f, int r) { return (T *)c->prog[f][r]; };
    {
        Timed _t(c, "advect_theta", (double)(ih - il) * (jh - jl) * (c->geo.NZ - 2));
        hipLaunchKernelGGL(k_advect_theta<T>, grd, blk, 0, c->stream, c->geo, c->cols, P(DYN_U, 1), P(DYN_V, 1), P(DYN_W, 1),
                       (const T *)c->bar[0], (const T *)c->bar[1], (const T *)c->bar[2], P(DYN_TH, 1), (const T *)c->bar[3],
                       P(DYN_TH, 0), P(DYN_TH, 2), T(step), il, jl, jh);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int advect_moisture_t(Ctx *c, double step, int il, int ih, int jl, int jh) {
    dim3 blk, grd = col_grid(c->geo, c->geo.NZ - 2, jl, jh, il, ih, blk);
    auto P = [&](int f, int r) { return (T *)c->prog[f][r]; };
    {
        Timed _t(c, "advect_moisture", (double)(ih - il) * (jh - jl) * (c->geo.NZ - 2));
        hipLaunchKernelGGL(k_advect_moisture<T>, grd, blk, 0, c->stream, c->geo, c->cols, P(DYN_U, 1), P(DYN_V, 1), P(DYN_W, 1),
                       (const T *)c->bar[0], (const T *)c->bar[1], (const T *)c->bar[2], P(DYN_QV, 1), P(DYN_QC, 1), P(DYN_QR, 1),
                       (const T *)c->bar[4], P(DYN_QV, 0), P(DYN_QC, 0), P(DYN_QR, 0), P(DYN_QV, 2), P(DYN_QC, 2), P(DYN_QR, 2),
                       (T *)c->vt, T(step), il, jl, jh);
    }
    DYN_LAUNCH_CHECK();
    // the stored fall speed is written after the advection kernel has read the old level NZ-1 values
    {
        Timed _t(c, "fall_speed", (double)(ih - il) * (jh - jl) * (c->geo.NZ - 2));
        hipLaunchKernelGGL(k_fall_speed<T>, grd, blk, 0, c->stream, c->geo, c->cols, P(DYN_QR, 1), (T *)c->vt, il, jl, jh);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int zero_moisture_t(Ctx *c, int il, int ih, int jl, int jh) {
    long long n = (long long)c->npts;
    {
        Timed _t(c, "zero_moisture", (double)n);
        hipLaunchKernelGGL(k_zero_moisture<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->geo, c->cols,
                       (T *)c->prog[DYN_QV][2], (T *)c->prog[DYN_QC][2], (T *)c->prog[DYN_QR][2], (const T *)c->bar[4], n, il, ih, jl, jh);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int pgf_t(Ctx *c, double step, int il, int ih, int jl, int jh) {
    dim3 blk, grd = col_grid(c->geo, c->geo.NZ - 2, jl, jh, il, ih, blk);
    {
        Timed _t(c, "pgf", (double)(ih - il) * (jh - jl) * (c->geo.NZ - 2));
        hipLaunchKernelGGL(k_pgf<T>, grd, blk, 0, c->stream, c->geo, c->cols, (const T *)c->pi, (T *)c->prog[DYN_U][2],
                       (T *)c->prog[DYN_V][2], (T *)c->prog[DYN_W][2], T(step), il, jl, jh);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int periodic_t(Ctx *c, void *const *fields, int nf) {
    FieldList fl{};
    fl.n = nf;
    for (int f = 0; f < nf; f++) fl.f[f] = fields[f];
    long long n = c->geo.NYNZ;
    {
        Timed _t(c, "periodic", 6.0 * n);
        hipLaunchKernelGGL(k_periodic<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->geo, fl);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int bc_t(Ctx *c, bool micro) {
    const Geo &g = c->geo;
    FieldList sp{}, per{}, ul{};
    int nfill;
    if (!micro) {
        void *UP = c->prog[DYN_U][2], *VP = c->prog[DYN_V][2], *WP = c->prog[DYN_W][2], *THP = c->prog[DYN_TH][2];
        sp = FieldList{{UP, VP, THP, WP}, 4};
        per = FieldList{{UP, VP, c->pi, THP, WP}, 5};
        ul = FieldList{{UP, VP, THP, WP}, 4};
        nfill = 3;
    } else {
        void *QV = c->prog[DYN_QV][2], *QC = c->prog[DYN_QC][2], *QR = c->prog[DYN_QR][2];
        sp = FieldList{{QV, QC, QR}, 3};
        per = sp;
        ul = sp;
        nfill = 3;
    }
    int rows = c->g.sponge_south + c->g.sponge_north;
    if (rows > 0) {
        {
        Timed _t(c, "bc_sponge", (double)rows * (g.NX - 1) * (g.NZ - 1));
        hipLaunchKernelGGL(k_sponge_ns<T>, dim3((g.NZ - 1 + 63) / 64, rows, g.NX - 1), dim3(64), 0, c->stream, g, sp, c->sponge_s,
                           c->sponge_n, c->g.sponge_south, c->g.sponge_north);
    }
        DYN_LAUNCH_CHECK();
    }
    long long n = g.NYNZ;
    {
        Timed _t(c, "bc_periodic", 6.0 * n);
        hipLaunchKernelGGL(k_periodic<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, g, per);
    }
    DYN_LAUNCH_CHECK();
    {
        Timed _t(c, "bc_upper_lower", 2.0 * (g.NX - 1) * (g.NY - 1));
        hipLaunchKernelGGL(k_upper_lower<T>, dim3((g.NY - 1 + 63) / 64, g.NX - 1), dim3(64), 0, c->stream, g, ul, nfill);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int kessler_t(Ctx *c, int il, int ih, int jl, int jh, KessTail tail = KessTail{}) {
    dim3 blk, grd = col_grid(c->geo, c->geo.NZ - 2, jl, jh, il, ih, blk);
    auto P = [&](int f) { return (T *)c->prog[f][2]; };
    {
        Timed _t(c, "kessler", (double)(ih - il) * (jh - jl) * (c->geo.NZ - 2));
        hipLaunchKernelGGL(k_kessler<T>, grd, blk, 0, c->stream, c->geo, c->cols, P(DYN_TH), P(DYN_QV), P(DYN_QC), P(DYN_QR),
                       (const T *)c->bar[3], (const T *)c->bar[4], (const T *)c->bar[5], (const T *)c->pi, c->dflag, c->turb.on ? c->turb.sat : nullptr, il,
                           jl, jh, tail);
    }
    DYN_LAUNCH_CHECK();
    {
        Timed _t(c, "precip", (double)(ih - il) * (jh - jl));
        hipLaunchKernelGGL(k_precip<T>, dim3((jh - jl + 63) / 64, ih - il), dim3(64), 0, c->stream, c->geo, c->cols, (const T *)c->vt,
                       P(DYN_QR), (T *)c->accrain, il, ih, jl, jh, tail);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int rayleigh_t(Ctx *c, int il, int ih, int jl, int jh, int level) {
    dim3 blk(64, 4, 1), grd((c->geo.NZ - level + 63) / 64, (jh - jl + 3) / 4, ih - il);
    auto P = [&](int f) { return (T *)c->prog[f][2]; };
    {
        Timed _t(c, "rayleigh", (double)(ih - il) * (jh - jl) * (c->geo.NZ - level));
        hipLaunchKernelGGL(k_rayleigh<T>, grd, blk, 0, c->stream, c->geo, c->raycoef, P(DYN_U), P(DYN_V), P(DYN_W), P(DYN_TH), il, jl, jh,
                       level);
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

}  // namespace

#define DISPATCH(c, fn, ...) ((c)->prec == DYN_FP32 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

int launch_advect_uvw(Ctx *c, double step, int il, int ih, int jl, int jh) { return DISPATCH(c, advect_uvw_t, c, step, il, ih, jl, jh); }
int launch_advect_theta(Ctx *c, double step, int il, int ih, int jl, int jh) { return DISPATCH(c, advect_theta_t, c, step, il, ih, jl, jh); }
int launch_advect_moisture(Ctx *c, double step, int il, int ih, int jl, int jh) {
    return DISPATCH(c, advect_moisture_t, c, step, il, ih, jl, jh);
}
int launch_zero_moisture(Ctx *c, int il, int ih, int jl, int jh) { return DISPATCH(c, zero_moisture_t, c, il, ih, jl, jh); }
int launch_pgf(Ctx *c, double step, int il, int ih, int jl, int jh) { return DISPATCH(c, pgf_t, c, step, il, ih, jl, jh); }
int launch_bc(Ctx *c, bool micro) { return DISPATCH(c, bc_t, c, micro); }
int launch_periodic(Ctx *c, void *const *fields, int nfields) { return DISPATCH(c, periodic_t, c, fields, nfields); }
int launch_kessler(Ctx *c, int il, int ih, int jl, int jh) { return DISPATCH(c, kessler_t, c, il, ih, jl, jh, KessTail{}); }
// Kessler over [1, NX-1) x [1, NY-1) with the lid and theta's damping (ray_level < 0: none) of the columns
// the step-end boundary update would change only at the lid: i in [6, NX-6), rows between the sponge
// and ghost rows
int launch_kessler_step_end(Ctx *c, int ray_level) {
    KessTail t;
    t.on = 1;
    t.i0 = 6;
    t.i1 = c->g.nx - 6;
    t.j0 = std::min(c->g.ny, std::max(c->g.sponge_south, 3));
    t.j1 = c->g.ny - std::min(c->g.ny - t.j0, std::max(c->g.sponge_north, 3));
    if (ray_level >= 0 && ray_level < c->g.nz) {
        t.lev = ray_level;
        t.ray = c->raycoef;
    }
    return DISPATCH(c, kessler_t, c, 1, c->g.nx - 1, 1, c->g.ny - 1, t);
}
int launch_rayleigh(Ctx *c, int il, int ih, int jl, int jh, int level) { return DISPATCH(c, rayleigh_t, c, il, ih, jl, jh, level); }

int launch_convert(Ctx *c, void *dst, const double *src, size_t n, bool to_device_type) {
    // FP32 contexts keep fields in float; the host side always speaks double
    long long nn = (long long)n;
    dim3 grd((unsigned)((nn + 255) / 256));
    if (to_device_type) hipLaunchKernelGGL(k_convert<float>, grd, dim3(256), 0, c->stream, (float *)dst, src, nn);
    else hipLaunchKernelGGL(k_convert_back<float>, grd, dim3(256), 0, c->stream, (double *)dst, (const float *)src, nn);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

}  // namespace dyn
