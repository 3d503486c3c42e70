// Turbulence parameterization (turbulence.cpp, USE_TURBULENT_STRESS) for the serial periodic path
// without terrain or surface heat flux: mixing coefficients and stored tendencies at the start of each
// step (calculate_diff_tend, driver_serial.cpp:194), added to the forecasts in every stage
// (apply_velocity_diffusion, solver.cpp:104; apply_scalar_diffusion / apply_moisture_diffusion,
// driver_serial.cpp:218, 231).
//
// The reference walks i-slabs and carries the stress tensor of slab i-1 over as the west stress of slab
// i, and the flux through the top of a level over as the bottom flux of the level above.  Every such
// value is a pure function of its position, so here one thread computes one point and re-evaluates the
// stress components and fluxes it needs; the values are those of the carried ones, in the reference's
// expression order (the library is built without FMA contraction).
#include <cmath>

#include "dyn_internal.hpp"

namespace dyn {
namespace {

constexpr double kCmixV = 0.15, kCmixH = 0.05;   // turbulence.cpp:48-49
constexpr double kPrInv = 3.0, kLvMax = 200.0, kCs = 0.25;  // turbulence.cpp:56-58

enum { TK = 0, TKH, TUF, TVF, TWF, TTD, TQV, TQC, TQR, TCOUNT };

struct TurbArgs {
    const double *vml2, *vkmax, *landsea;
    double kmixh_max, lH;
};

// the fields the coefficients and tendencies read: the M buffers, the basic state, pi and the
// saturation flags Kessler leaves
template <typename T>
struct TurbIn {
    const T *um, *vm, *wm, *thm, *qvm, *qcm, *qrm;
    const T *ub, *vb, *wb, *thb, *qb, *pb, *pi;
    const unsigned char *sat;
    int moist;
};

#define IX3(i, j, k) ((long long)(i) * g.NYNZ + (long long)(j) * g.NZ + (k))

template <typename T>
__device__ __forceinline__ T odz(const Geo &g, const Cols &c, int k) { return T(g.one_d_dz) * T(c.mu[k]); }
template <typename T>
__device__ __forceinline__ T odzw(const Geo &g, const Cols &c, int k) { return T(g.one_d_dz) * T(c.mw[k]); }

// S^2 = D11^2 + ... + D23^2, turbulence.cpp:596-624
template <typename T>
__device__ T deformation2(const Geo &g, const Cols &c, const TurbIn<T> &f, int i, int j, int k) {
    const T *UM = f.um, *VM = f.vm, *WM = f.wm, *UB = f.ub, *VB = f.vb, *WB = f.wb;
    const T odx = T(g.one_d_dx), ody = T(g.one_d_dy), dzw = odzw<T>(g, c, k);
#define A(F, a, b, cc) F[IX3(a, b, cc)]
    T D11 = T(0.50) * (A(UM, i + 1, j, k) + A(UM, i + 1, j, k - 1) - A(UM, i - 1, j, k) - A(UM, i - 1, j, k - 1) +
                       A(UB, i + 1, j, k) + A(UB, i + 1, j, k - 1) - A(UB, i - 1, j, k) - A(UB, i - 1, j, k - 1)) * odx;
    T D22 = T(0.50) * (A(VM, i, j + 1, k) + A(VM, i, j + 1, k - 1) - A(VM, i, j - 1, k) - A(VM, i, j - 1, k - 1) +
                       A(VB, i, j + 1, k) + A(VB, i, j + 1, k - 1) - A(VB, i, j - 1, k) - A(VB, i, j - 1, k - 1)) * ody;
    T D33 = T(1.0) * (A(WM, i, j, k + 1) - A(WM, i, j, k - 1) + A(WB, i, j, k + 1) - A(WB, i, j, k - 1)) * dzw;
    T D12 = T(0.5) * ((A(UM, i, j + 1, k) + A(UM, i, j + 1, k - 1) - A(UM, i, j - 1, k) - A(UM, i, j - 1, k - 1) +
                       A(UB, i, j + 1, k) + A(UB, i, j + 1, k - 1) - A(UB, i, j - 1, k) - A(UB, i, j - 1, k - 1)) * ody +
                      (A(VM, i + 1, j, k) + A(VM, i + 1, j, k - 1) - A(VM, i - 1, j, k) - A(VM, i - 1, j, k - 1) +
                       A(VB, i + 1, j, k) + A(VB, i + 1, j, k - 1) - A(VB, i - 1, j, k) - A(VB, i - 1, j, k - 1)) * odx);
    T D13 = T(0.5) * ((A(UM, i, j, k) + A(UM, i + 1, j, k) - A(UM, i, j, k - 1) - A(UM, i + 1, j, k - 1) +
                       A(UB, i, j, k) + A(UB, i + 1, j, k) - A(UB, i, j, k - 1) - A(UB, i + 1, j, k - 1)) * dzw +
                      (A(WM, i + 1, j, k) - A(WM, i - 1, j, k) + A(WB, i + 1, j, k) - A(WB, i - 1, j, k)) * odx);
    T D23 = T(0.5) * ((A(VM, i, j, k) + A(VM, i, j + 1, k) - A(VM, i, j, k - 1) - A(VM, i, j + 1, k - 1) +
                       A(VB, i, j, k) + A(VB, i, j + 1, k) - A(VB, i, j, k - 1) - A(VB, i, j + 1, k - 1)) * dzw +
                      (A(WM, i, j + 1, k) - A(WM, i, j - 1, k) + A(WB, i, j + 1, k) - A(WB, i, j - 1, k)) * ody);
#undef A
    return D11 * D11 + D22 * D22 + D33 * D33 + D12 * D12 + D13 * D13 + D23 * D23;
}

// Brunt-Vaisala frequency: get_Kmix_dry (turbulence.cpp:781-791), get_Kmix (:627-690)
template <typename T>
__device__ T brunt_vaisala(const Geo &g, const Cols &c, const TurbIn<T> &f, int i, int j, int k) {
    const long long h = IX3(i, j, k), l = IX3(i, j, k - 1);
    if (!f.moist || !f.sat[h] || !f.sat[l]) {
        T qh = f.moist ? f.qvm[h] + f.qb[h] + T(c.qb[k]) : T(c.qb[k]);
        T ql = f.moist ? f.qvm[l] + f.qb[l] + T(c.qb[k - 1]) : T(c.qb[k - 1]);
        T thetaVH = (f.thm[h] + f.thb[h] + T(c.tb[k])) * (T(1) + T(0.61) * qh);
        T thetaVL = (f.thm[l] + f.thb[l] + T(c.tb[k - 1])) * (T(1) + T(0.61) * ql);
        T dTheta = thetaVH - thetaVL;
        return T(2.0) * T(kGrav) / (T(c.tbv[k]) + T(c.tbv[k - 1])) * dTheta * odzw<T>(g, c, k);
    }
    const T cp = T(kCp), Lv = T(kLv), Rd = T(kRd), grav = T(kGrav);
    const T cpRd = T(kCp / kRd);
    T thetaVH = f.thm[h] + f.thb[h] + T(c.tb[k]);
    T thetaVL = f.thm[l] + f.thb[l] + T(c.tb[k - 1]);
    T tempH = thetaVH * (f.pb[h] + f.pi[h] / (cp * T(c.tbv[k])));
    T tempL = thetaVL * (f.pb[l] + f.pi[l] / (cp * T(c.tbv[k - 1])));
    T temp = T(0.5) * (tempH + tempL);
    T dQL = f.qcm[h] + f.qrm[h] - f.qcm[l] - f.qrm[l] + f.qvm[h] - f.qvm[l] + f.qb[h] - f.qb[l] + T(c.qb[k]) - T(c.qb[k - 1]);
    T qv = T(0.5) * (f.qvm[h] + f.qvm[l] + f.qb[h] + f.qb[l] + T(c.qb[k]) + T(c.qb[k - 1]));
    T pdH = T(kP0) * pow(f.pb[h], cpRd) + f.pi[h] * T(c.rhou[k]);
    T pdL = T(kP0) * pow(f.pb[l], cpRd) + f.pi[l] * T(c.rhou[k - 1]);
    T eslH = T(611.2) * exp(T(17.67) * (tempH - T(273.15)) / (tempH - T(29.65)));
    T eslL = T(611.2) * exp(T(17.67) * (tempL - T(273.15)) / (tempL - T(29.65)));
    T qvsatH = T(0.62197) * eslH / (pdH - eslH);
    T qvsatL = T(0.62197) * eslL / (pdL - eslL);
    thetaVH = thetaVH * (T(1) + Lv * qvsatH / (cp * tempH));
    thetaVL = thetaVL * (T(1) + Lv * qvsatL / (cp * tempL));
    T dTheta = thetaVH - thetaVL;
    T Acoef = (grav / T(c.tbw[k])) * (T(1.0) + Lv * qv / (Rd * temp)) / (T(1.0) + T(0.622) * Lv * Lv * qv / (cp * Rd * temp * temp));
    return (Acoef * dTheta - grav * dQL) * odzw<T>(g, c, k);
}

// the vertical coefficient of level k before (raw) and after (final) its bound; levels 1 and NZ-1 are 0
template <typename T>
__device__ void kmix_level(const Geo &g, const Cols &c, const TurbIn<T> &f, const TurbArgs &ta, int i, int j, int k, T *raw,
                           T *final_) {
    *raw = T(0);
    *final_ = T(0);
    if (k < 2 || k > g.NZ - 2) return;
    T S2 = deformation2<T>(g, c, f, i, j, k);
    T bruntv = brunt_vaisala<T>(g, c, f, i, j, k);
    if (S2 - T(kPrInv) * bruntv > T(0)) {
        T cc = T(ta.vml2[k]) * T(kCs) * T(kCs);
        *raw = cc * sqrt(S2 - T(kPrInv) * bruntv);
        *final_ = *raw > T(ta.vkmax[k]) ? T(ta.vkmax[k]) : *raw;
    }
}

// get_Kmix / get_Kmix_dry, turbulence.cpp:563-831: one thread per (i, j, k), k in [1, NZ); the
// horizontal coefficient of level k-1 is set with level k, from level k-1's final and level k's raw value
template <typename T>
__global__ void __launch_bounds__(256) k_turb_kmix(Geo g, Cols c, TurbIn<T> f, TurbArgs ta, T *__restrict__ K, T *__restrict__ KH,
                                                   int il, int jl, int jh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + jl, i = blockIdx.z + il;
    if (k >= g.NZ || j >= jh) return;
    const long long p = IX3(i, j, k);
    if (k == 1 || k == g.NZ - 1) {
        K[p] = T(0);
        return;
    }
    T raw, fin, rawl, finl;
    kmix_level<T>(g, c, f, ta, i, j, k, &raw, &fin);
    kmix_level<T>(g, c, f, ta, i, j, k - 1, &rawl, &finl);
    const T lH = T(ta.lH);
    T kh;
    if (raw > T(0)) {
        kh = T(0.5) * (finl / T(ta.vml2[k - 1]) + raw / T(ta.vml2[k])) * lH * lH;
        if (kh > T(ta.kmixh_max)) kh = T(ta.kmixh_max);
    } else {
        kh = T(0) + T(0.5) * finl / T(ta.vml2[k - 1]) * lH * lH;
    }
    K[p] = fin;
    KH[p - 1] = kh;
}

// stress components of slab i (turbulence.cpp:913-933); level 0 of tau_33, tau_13, tau_23 is zero
template <typename T>
__device__ __forceinline__ T tau11(const Geo &g, const TurbIn<T> &f, int i, int j, int k) {
    return T(2.0) * (f.um[IX3(i + 1, j, k)] - f.um[IX3(i, j, k)]) * T(g.one_d_dx);
}
template <typename T>
__device__ __forceinline__ T tau22(const Geo &g, const TurbIn<T> &f, int i, int j, int k) {
    return T(2.0) * (f.vm[IX3(i, j + 1, k)] - f.vm[IX3(i, j, k)]) * T(g.one_d_dy);
}
template <typename T>
__device__ __forceinline__ T tau33(const Geo &g, const Cols &c, const TurbIn<T> &f, int i, int j, int k) {
    return k < 1 ? T(0) : T(2.0) * (f.wm[IX3(i, j, k + 1)] - f.wm[IX3(i, j, k)]) * odz<T>(g, c, k);
}
template <typename T>
__device__ __forceinline__ T tau13(const Geo &g, const Cols &c, const TurbIn<T> &f, int i, int j, int k) {
    if (k < 1) return T(0);
    return (f.wm[IX3(i + 1, j, k + 1)] - f.wm[IX3(i, j, k + 1)]) * T(g.one_d_dx) +
           (f.um[IX3(i + 1, j, k + 1)] - f.um[IX3(i + 1, j, k)]) * odzw<T>(g, c, k + 1);
}
template <typename T>
__device__ __forceinline__ T tau23(const Geo &g, const Cols &c, const TurbIn<T> &f, int i, int j, int k) {
    if (k < 1) return T(0);
    return (f.wm[IX3(i, j + 1, k + 1)] - f.wm[IX3(i, j, k + 1)]) * T(g.one_d_dy) +
           (f.vm[IX3(i, j + 1, k + 1)] - f.vm[IX3(i, j + 1, k)]) * odzw<T>(g, c, k + 1);
}
template <typename T>
__device__ __forceinline__ T tau12(const Geo &g, const TurbIn<T> &f, int i, int j, int k) {
    return (f.um[IX3(i + 1, j + 1, k)] - f.um[IX3(i + 1, j, k)]) * T(g.one_d_dy) +
           (f.vm[IX3(i + 1, j + 1, k)] - f.vm[IX3(i, j + 1, k)]) * T(g.one_d_dx);
}

// turbulent_diffusion_velocity, turbulence.cpp:839-1020: friction tendencies of u, v, w at (i, j, k),
// k in [1, NZ-1); the bottom flux of level 1 is the bulk surface stress (turbulence.cpp:885-907)
template <typename T>
__global__ void __launch_bounds__(256) k_turb_velocity(Geo g, Cols c, TurbIn<T> f, TurbArgs ta, const T *__restrict__ K,
                                                       const T *__restrict__ KH, T *__restrict__ UF, T *__restrict__ VF,
                                                       T *__restrict__ WF, int il, int jl, int jh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + jl, i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    const T odx = T(g.one_d_dx), ody = T(g.one_d_dy);
#define KM(a, b, cc) K[IX3(a, b, cc)]
#define KHM(a, b, cc) KH[IX3(a, b, cc)]
    T tau_u_zl, tau_v_zl;
    if (k == 1) {
        const long long n = IX3(i, j, 1);
        T wind_speed = sqrt((f.um[n] + f.ub[n]) * (f.um[n] + f.ub[n]) + (f.vm[n] + f.vb[n]) * (f.vm[n] + f.vb[n]));
        T wind_speed_base = sqrt(f.ub[n] * f.ub[n] + f.vb[n] * f.vb[n]);
        if (ta.landsea[(long long)i * g.NY + j] < 0.5) {
            tau_u_zl = T(0.003) * (wind_speed * f.um[n] + (wind_speed - wind_speed_base) * f.ub[n]);
            tau_v_zl = T(0.003) * (wind_speed * f.vm[n] + (wind_speed - wind_speed_base) * f.vb[n]);
        } else {
            T drag_coef = (T(1.1e-3) + T(4.0e-5) * wind_speed);
            tau_u_zl = drag_coef * (wind_speed * f.um[n] + (wind_speed - wind_speed_base) * f.ub[n]);
            tau_v_zl = drag_coef * (wind_speed * f.vm[n] + (wind_speed - wind_speed_base) * f.vb[n]);
        }
    } else {  // the top flux of level k-1
        T uK13V_l = T(0.50) * (KM(i, j, k) + KM(i - 1, j, k));
        T vK23V_l = T(0.5) * (KM(i, j, k) + KM(i, j - 1, k));
        tau_u_zl = uK13V_l * tau13<T>(g, c, f, i - 1, j, k - 1);
        tau_v_zl = vK23V_l * tau23<T>(g, c, f, i, j - 1, k - 1);
    }
    T uK11H_h = KHM(i, j, k);
    T uK11H_l = KHM(i - 1, j, k);
    T uK12H_h = T(0.25) * (KHM(i, j, k) + KHM(i - 1, j, k) + KHM(i - 1, j + 1, k) + KHM(i, j + 1, k));
    T uK12H_l = T(0.25) * (KHM(i, j - 1, k) + KHM(i - 1, j - 1, k) + KHM(i - 1, j, k) + KHM(i, j, k));
    T uK13V_h = T(0.50) * (KM(i, j, k + 1) + KM(i - 1, j, k + 1));
    T vK12H_h = T(0.25) * (KHM(i, j - 1, k) + KHM(i, j, k) + KHM(i + 1, j - 1, k) + KHM(i + 1, j, k));
    T vK12H_l = uK12H_l;
    T vK22H_h = KHM(i, j, k);
    T vK22H_l = KHM(i, j - 1, k);
    T vK23V_h = T(0.5) * (KM(i, j, k + 1) + KM(i, j - 1, k + 1));
    T wK13H_h = T(0.25) * (KHM(i, j, k) + KHM(i, j, k - 1) + KHM(i + 1, j, k) + KHM(i + 1, j, k - 1));
    T wK13H_l = T(0.25) * (KHM(i, j, k) + KHM(i, j, k - 1) + KHM(i - 1, j, k) + KHM(i - 1, j, k - 1));
    T wK23H_h = T(0.25) * (KHM(i, j, k) + KHM(i, j, k - 1) + KHM(i, j + 1, k) + KHM(i, j + 1, k - 1));
    T wK23H_l = T(0.25) * (KHM(i, j, k) + KHM(i, j, k - 1) + KHM(i, j - 1, k) + KHM(i, j - 1, k - 1));
    T wK33V_h = T(0.50) * (KM(i, j, k) + KM(i, j, k + 1));
    T wK33V_l = T(0.50) * (KM(i, j, k) + KM(i, j, k - 1));
#undef KM
#undef KHM
    T ufric = (uK11H_h * tau11<T>(g, f, i, j, k) - uK11H_l * tau11<T>(g, f, i - 1, j, k)) * odx +
              (uK12H_h * tau12<T>(g, f, i - 1, j, k) - uK12H_l * tau12<T>(g, f, i - 1, j - 1, k)) * ody +
              (uK13V_h * tau13<T>(g, c, f, i - 1, j, k) - tau_u_zl) * odz<T>(g, c, k);
    T vfric = (vK12H_h * tau12<T>(g, f, i, j - 1, k) - vK12H_l * tau12<T>(g, f, i - 1, j - 1, k)) * odx +
              (vK22H_h * tau22<T>(g, f, i, j, k) - vK22H_l * tau22<T>(g, f, i, j - 1, k)) * ody +
              (vK23V_h * tau23<T>(g, c, f, i, j - 1, k) - tau_v_zl) * odz<T>(g, c, k);
    T wfric = (wK13H_h * tau13<T>(g, c, f, i, j, k - 1) - wK13H_l * tau13<T>(g, c, f, i - 1, j, k - 1)) * odx +
              (wK23H_h * tau23<T>(g, c, f, i, j, k - 1) - wK23H_l * tau23<T>(g, c, f, i, j - 1, k - 1)) * ody +
              (wK33V_h * tau33<T>(g, c, f, i, j, k) - wK33V_l * tau33<T>(g, c, f, i, j, k - 1)) * odzw<T>(g, c, k);
    const long long p = IX3(i, j, k);
    UF[p] = ufric;
    VF[p] = vfric;
    WF[p] = wfric;
}

// the top diffusive flux of level k of scalar S, turbulence.cpp:1128-1131, 1226
template <typename T>
__device__ __forceinline__ T scalar_top_flux(const Geo &g, const Cols &c, const T *K, const T *S, int i, int j, int k) {
    return -T(kPrInv) * K[IX3(i, j, k + 1)] * (S[IX3(i, j, k + 1)] - S[IX3(i, j, k)]) * odzw<T>(g, c, k + 1);
}

// turbulent_diffusion_theta / turbulent_diffusion_scalars, turbulence.cpp:1027-1276: theta (and qv, qc,
// qr) at (i, j, k), k in [1, NZ-1); no flux through the ground
template <typename T>
__global__ void __launch_bounds__(256) k_turb_scalars(Geo g, Cols c, TurbIn<T> f, const T *__restrict__ K, const T *__restrict__ KH,
                                                      T *__restrict__ TD, T *__restrict__ QVD, T *__restrict__ QCD, T *__restrict__ QRD,
                                                      int il, int jl, int jh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + jl, i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    const T Pr = T(kPrInv), odx = T(g.one_d_dx), ody = T(g.one_d_dy);
    const T *src[4] = {f.thm, f.qvm, f.qcm, f.qrm};
    T *dst[4] = {TD, QVD, QCD, QRD};
    const int nf = f.moist ? 4 : 1;
    const long long p = IX3(i, j, k);
    // the horizontal coefficients, shared by every scalar
    const T KHn = T(0.5) * (Pr * KH[IX3(i, j + 1, k)] + Pr * KH[p]);
    const T KHs = T(0.5) * (Pr * KH[p] + Pr * KH[IX3(i, j - 1, k)]);
    const T KHe = T(0.5) * (Pr * KH[IX3(i + 1, j, k)] + Pr * KH[p]);
    const T KHw = T(0.5) * (Pr * KH[p] + Pr * KH[IX3(i - 1, j, k)]);
    for (int q = 0; q < nf; q++) {
        const T *S = src[q];
        T tau_h = scalar_top_flux<T>(g, c, K, S, i, j, k);
        T tau_l = k == 1 ? T(0) : scalar_top_flux<T>(g, c, K, S, i, j, k - 1);
        T t22h = -KHn * (S[IX3(i, j + 1, k)] - S[p]) * ody;
        T t22l = -KHs * (S[p] - S[IX3(i, j - 1, k)]) * ody;
        T t11h = -KHe * (S[IX3(i + 1, j, k)] - S[p]) * odx;
        T t11l = -KHw * (S[p] - S[IX3(i - 1, j, k)]) * odx;
        dst[q][p] = -((t11h - t11l) * odx + (t22h - t22l) * ody + (tau_h - tau_l) * odz<T>(g, c, k));
    }
}

// apply_velocity_diffusion, turbulence.cpp:233-258 (w from level 2)
template <typename T>
__global__ void k_turb_apply_uvw(Geo g, const T *__restrict__ UF, const T *__restrict__ VF, const T *__restrict__ WF, T *__restrict__ UP,
                                 T *__restrict__ VP, T *__restrict__ WP, T deltaT, int il, int jl, int jh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + jl, i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    const long long p = IX3(i, j, k);
    UP[p] += UF[p] * deltaT;
    VP[p] += VF[p] * deltaT;
    if (k >= 2) WP[p] += WF[p] * deltaT;
}

// apply_scalar_diffusion (THP += t * step * dt) and apply_moisture_diffusion (QP += q * (step*dt)),
// turbulence.cpp:263-292
template <typename T>
__global__ void k_turb_apply_scalars(Geo g, const T *__restrict__ D0, const T *__restrict__ D1, const T *__restrict__ D2,
                                     T *__restrict__ P0, T *__restrict__ P1, T *__restrict__ P2, int nf, int theta, T step, T dt,
                                     int il, int jl, int jh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + jl, i = blockIdx.z + il;
    if (k >= g.NZ - 1 || j >= jh) return;
    const long long p = IX3(i, j, k);
    if (theta) {
        P0[p] += D0[p] * step * dt;
        return;
    }
    const T deltaT = step * dt;
    P0[p] += D0[p] * deltaT;
    if (nf > 1) P1[p] += D1[p] * deltaT;
    if (nf > 2) P2[p] += D2[p] * deltaT;
}

#undef IX3

template <typename T>
TurbIn<T> turb_inputs(Ctx *c) {
    TurbIn<T> f{};
    auto M = [&](int fld) { return (const T *)c->prog[fld][DYN_ROLE_M]; };
    f.um = M(DYN_U);
    f.vm = M(DYN_V);
    f.wm = M(DYN_W);
    f.thm = M(DYN_TH);
    f.moist = c->g.microphysics != 0;
    if (f.moist) {
        f.qvm = M(DYN_QV);
        f.qcm = M(DYN_QC);
        f.qrm = M(DYN_QR);
    }
    f.ub = (const T *)c->bar[0];
    f.vb = (const T *)c->bar[1];
    f.wb = (const T *)c->bar[2];
    f.thb = (const T *)c->bar[3];
    f.qb = (const T *)c->bar[4];
    f.pb = (const T *)c->bar[5];
    f.pi = (const T *)c->pi;
    f.sat = c->turb.sat;
    return f;
}

TurbArgs turb_args(const Ctx *c) {
    return TurbArgs{c->turb.vml2, c->turb.vkmax, c->turb.landsea, c->turb.kmixh_max, c->turb.lH};
}

dim3 grid3(const Ctx *c, int il, int ih, int jl, int jh, int kblock) {
    return dim3((c->g.nz + kblock - 1) / kblock, (unsigned)(jh - jl), (unsigned)(ih - il));
}

template <typename T>
int launch_diff_tend(Ctx *c, int il, int ih, int jl, int jh) {
    TurbIn<T> f = turb_inputs<T>(c);
    TurbArgs ta = turb_args(c);
    T **b = (T **)c->turb.buf;
    const int kb = 64;
    hipLaunchKernelGGL(k_turb_kmix<T>, grid3(c, il, ih, jl, jh, kb), dim3(kb), 0, c->stream, c->geo, c->cols, f, ta, b[TK], b[TKH], il, jl,
                       jh);
    DYN_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_turb_velocity<T>, grid3(c, il, ih, jl, jh, kb), dim3(kb), 0, c->stream, c->geo, c->cols, f, ta, (const T *)b[TK],
                       (const T *)b[TKH], b[TUF], b[TVF], b[TWF], il, jl, jh);
    DYN_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_turb_scalars<T>, grid3(c, il, ih, jl, jh, kb), dim3(kb), 0, c->stream, c->geo, c->cols, f, (const T *)b[TK],
                       (const T *)b[TKH], b[TTD], b[TQV], b[TQC], b[TQR], il, jl, jh);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int launch_apply_velocity(Ctx *c, double step, int il, int ih, int jl, int jh) {
    T **b = (T **)c->turb.buf;
    auto P = [&](int fld) { return (T *)c->prog[fld][DYN_ROLE_P]; };
    const double deltaT = step * c->g.dt;
    hipLaunchKernelGGL(k_turb_apply_uvw<T>, grid3(c, il, ih, jl, jh, 64), dim3(64), 0, c->stream, c->geo, (const T *)b[TUF],
                       (const T *)b[TVF], (const T *)b[TWF], P(DYN_U), P(DYN_V), P(DYN_W), T(deltaT), il, jl, jh);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int launch_apply_scalars(Ctx *c, double step, int il, int ih, int jl, int jh, bool theta) {
    T **b = (T **)c->turb.buf;
    auto P = [&](int fld) { return (T *)c->prog[fld][DYN_ROLE_P]; };
    if (theta)
        hipLaunchKernelGGL(k_turb_apply_scalars<T>, grid3(c, il, ih, jl, jh, 64), dim3(64), 0, c->stream, c->geo, (const T *)b[TTD],
                           (const T *)nullptr, (const T *)nullptr, P(DYN_TH), (T *)nullptr, (T *)nullptr, 1, 1, T(step), T(c->g.dt), il,
                           jl, jh);
    else
        hipLaunchKernelGGL(k_turb_apply_scalars<T>, grid3(c, il, ih, jl, jh, 64), dim3(64), 0, c->stream, c->geo, (const T *)b[TQV],
                           (const T *)b[TQC], (const T *)b[TQR], P(DYN_QV), P(DYN_QC), P(DYN_QR), 3, 0, T(step), T(c->g.dt), il, jl,
                           jh);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

int check_turb(Ctx *c, const char *who, int il, int ih, int jl, int jh, int lo_i, int hi_i, int lo_j, int hi_j) {
    if (!c) {
        set_error(std::string(who) + ": null context");
        return DYN_E_ARG;
    }
    if (!c->turb.on) {
        set_error(std::string(who) + ": turbulence is off (dyn_set_turbulence)");
        return DYN_E_STATE;
    }
    if (il < lo_i || ih > hi_i || jl < lo_j || jh > hi_j || il >= ih || jl >= jh) {
        set_error(std::string(who) + ": bounds outside the range the reference uses");
        return DYN_E_ARG;
    }
    return DYN_OK;
}

}  // namespace

void turb_release(Ctx *c) {
    for (void *&p : c->turb.buf)
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
    void *ptrs[] = {c->turb.sat, c->turb.landsea, c->turb.vml2, c->turb.vkmax};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c->turb = Turbulence{};
}

int turb_step_begin(Ctx *c) {
    const int NX = c->g.nx, NY = c->g.ny;
    return c->prec == DYN_FP32 ? launch_diff_tend<float>(c, 1, NX - 1, 1, NY - 1) : launch_diff_tend<double>(c, 1, NX - 1, 1, NY - 1);
}
int turb_apply_velocity(Ctx *c, double step) {
    const int NX = c->g.nx, NY = c->g.ny;
    return c->prec == DYN_FP32 ? launch_apply_velocity<float>(c, step, 3, NX - 3, 3, NY - 3)
                               : launch_apply_velocity<double>(c, step, 3, NX - 3, 3, NY - 3);
}
int turb_apply_scalars(Ctx *c, double step, bool theta) {
    const int NX = c->g.nx, NY = c->g.ny;
    return c->prec == DYN_FP32 ? launch_apply_scalars<float>(c, step, 1, NX - 1, 1, NY - 1, theta)
                               : launch_apply_scalars<double>(c, step, 1, NX - 1, 1, NY - 1, theta);
}

}  // namespace dyn

using namespace dyn;

extern "C" {

int dyn_set_turbulence(dyn_ctx *ctx, int on, const double *landsea, size_t count) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) {
        set_error("dyn_set_turbulence: null context");
        return DYN_E_ARG;
    }
    const size_t n2 = (size_t)c->g.nx * c->g.ny;
    if (landsea && count != n2) {
        set_error("dyn_set_turbulence: the land-sea mask needs NX*NY values");
        return DYN_E_ARG;
    }
    if (int rc = dyn_synchronize(ctx)) return rc;
    drop_graph(c);  // the step's kernel sequence changes
    if (!on) {
        turb_release(c);
        return DYN_OK;
    }
    Turbulence &t = c->turb;
    if (!t.buf[0]) {
        for (void *&p : t.buf) {
            if (hipMalloc(&p, c->npts * c->elt) != hipSuccess || hipMemset(p, 0, c->npts * c->elt) != hipSuccess) {
                turb_release(c);
                set_error("dyn_set_turbulence: out of device memory");
                return DYN_E_HIP;
            }
        }
        if (hipMalloc(&t.sat, c->npts) != hipSuccess || hipMemset(t.sat, 0, c->npts) != hipSuccess ||
            hipMalloc(&t.landsea, n2 * 8) != hipSuccess || hipMalloc(&t.vml2, (size_t)c->g.nz * 8) != hipSuccess ||
            hipMalloc(&t.vkmax, (size_t)c->g.nz * 8) != hipSuccess) {
            turb_release(c);
            set_error("dyn_set_turbulence: out of device memory");
            return DYN_E_HIP;
        }
    }
    // init_kmix, turbulence.cpp:85-147, with zlevs = ZU = zsu (driver_serial.cpp:52)
    const int NZ = c->g.nz;
    const double *zsu = c->colhost + (size_t)DYN_COL_ZSU * NZ;
    std::vector<double> vml2(NZ, 0.0), vkmax(NZ, 0.0), ls(n2, 1.0);
    for (int k = 1; k < NZ; k++) {
        double deltaZ = zsu[k] - zsu[k - 1];
        vml2[k] = deltaZ > kLvMax ? kLvMax * kLvMax : deltaZ * deltaZ;
        vkmax[k] = kCmixV * vml2[k] / c->g.dt;
    }
    if (landsea)
        for (size_t n = 0; n < n2; n++) ls[n] = landsea[n];
    t.kmixh_max = kCmixH * c->g.dx * c->g.dx / c->g.dt;
    t.lH = c->g.dx / 4.0;
    DYN_HIP(hipMemcpy(t.vml2, vml2.data(), NZ * 8, hipMemcpyHostToDevice));
    DYN_HIP(hipMemcpy(t.vkmax, vkmax.data(), NZ * 8, hipMemcpyHostToDevice));
    DYN_HIP(hipMemcpy(t.landsea, ls.data(), n2 * 8, hipMemcpyHostToDevice));
    t.on = true;
    return DYN_OK;
}

int dyn_calculate_diff_tend(dyn_ctx *ctx, int il, int ih, int jl, int jh) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_turb(c, "calculate_diff_tend", il, ih, jl, jh, 1, c ? c->g.nx - 1 : 0, 1, c ? c->g.ny - 1 : 0)) return rc;
    return c->prec == DYN_FP32 ? launch_diff_tend<float>(c, il, ih, jl, jh) : launch_diff_tend<double>(c, il, ih, jl, jh);
}

int dyn_apply_velocity_diffusion(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_turb(c, "apply_velocity_diffusion", il, ih, jl, jh, 1, c ? c->g.nx - 1 : 0, 1, c ? c->g.ny - 1 : 0)) return rc;
    return c->prec == DYN_FP32 ? launch_apply_velocity<float>(c, step, il, ih, jl, jh) : launch_apply_velocity<double>(c, step, il, ih, jl, jh);
}

int dyn_apply_scalar_diffusion(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_turb(c, "apply_scalar_diffusion", il, ih, jl, jh, 1, c ? c->g.nx - 1 : 0, 1, c ? c->g.ny - 1 : 0)) return rc;
    return c->prec == DYN_FP32 ? launch_apply_scalars<float>(c, step, il, ih, jl, jh, true)
                               : launch_apply_scalars<double>(c, step, il, ih, jl, jh, true);
}

int dyn_apply_moisture_diffusion(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_turb(c, "apply_moisture_diffusion", il, ih, jl, jh, 1, c ? c->g.nx - 1 : 0, 1, c ? c->g.ny - 1 : 0)) return rc;
    if (!c->g.microphysics) {
        set_error("apply_moisture_diffusion: context has no moisture fields");
        return DYN_E_STATE;
    }
    return c->prec == DYN_FP32 ? launch_apply_scalars<float>(c, step, il, ih, jl, jh, false)
                               : launch_apply_scalars<double>(c, step, il, ih, jl, jh, false);
}

int dyn_get_turbulence(dyn_ctx *ctx, int which, double *values, size_t count) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !c->turb.on || which < 0 || which > TCOUNT || !values || count != c->npts) {
        set_error("dyn_get_turbulence: turbulence off, bad array id or count other than NX*NY*NZ");
        return DYN_E_ARG;
    }
    DYN_HIP(hipStreamSynchronize(c->stream));
    if (which == TCOUNT) {  // the saturation flags
        std::vector<unsigned char> tmp(c->npts);
        DYN_HIP(hipMemcpy(tmp.data(), c->turb.sat, c->npts, hipMemcpyDeviceToHost));
        for (size_t n = 0; n < c->npts; n++) values[n] = tmp[n];
        return DYN_OK;
    }
    if (c->prec == DYN_FP64) {
        DYN_HIP(hipMemcpy(values, c->turb.buf[which], c->npts * 8, hipMemcpyDeviceToHost));
        return DYN_OK;
    }
    std::vector<float> tmp(c->npts);
    DYN_HIP(hipMemcpy(tmp.data(), c->turb.buf[which], c->npts * 4, hipMemcpyDeviceToHost));
    for (size_t n = 0; n < c->npts; n++) values[n] = tmp[n];
    return DYN_OK;
}

}  // extern "C"
