// Hand-written discrete adjoint of the dry step (s_integrate_rk3 with MICROPHYSICS_OPTION 0,
// driver_serial.cpp:178-286): advection and flux stencils (advection.cpp:304-373, fluxes.cpp), the
// pressure projection (solve_pressure + the gradient force of integrate_non_hydro, solver.cpp:87-138),
// boundary conditions (boundaries.cpp:119-193), Rayleigh damping (damping.cpp:1148-1177) and the
// three-buffer rotation (solver.cpp:229-263).
//
// The adjoint state mirrors the forward's buffers: every prognostic keeps adjoints of its M, current and
// forecast buffers, rotated in reverse of the forward's pointer swaps, because the forward reads stale
// ghost rows of rotated buffers.  Upwind signs are held fixed (the step is piecewise linear in the
// perturbation fields at fixed signs); everything else is the exact transpose of the forward arithmetic.
//
// Trajectory: a forward pass stores the 12 buffers (U V W TH x M, current, forecast) every `every` steps;
// the reverse sweep restores each snapshot, recomputes its segment while recording the current buffers at
// the start of every RK stage, and runs the adjoint steps of the segment backwards (two-level
// checkpointing).
#include <cstring>
#include <string>
#include <vector>

#include "dyn_internal.hpp"

namespace dyn {
namespace {

#define IX(i, j, k) ((long long)(i) * g.NYNZ + (long long)(j) * g.NZ + (long long)(k))

constexpr int kAdjFields = 4;  // U V W TH
const int kAdjIds[kAdjFields] = {DYN_U, DYN_V, DYN_W, DYN_TH};

template <typename T>
__device__ __forceinline__ T sgn(T x) { return x > T(0) ? T(1) : T(-1); }
template <typename T>
__device__ __forceinline__ T i5(T pm2, T pm1, T p0, T p1, T p2, T p3, T sign) {
    const T c37_60 = T(37.0 / 60.0), c2_15 = T(2.0 / 15.0), c1_60 = T(1.0 / 60.0);
    T c6 = c37_60 * (p1 + p0) - c2_15 * (p2 + pm1) + c1_60 * (p3 + pm2);
    T d5 = -(p3 - pm2) + T(5.0) * (p2 - pm1) - T(10.0) * (p1 - p0);
    return c6 + c1_60 * sign * d5;
}
template <typename T>
__device__ __forceinline__ T i3(T pm1, T p0, T p1, T p2, T sign) {
    const T c7_12 = T(7.0 / 12.0), c1_12 = T(1.0 / 12.0);
    return (c7_12 * (p1 + p0) - c1_12 * (p2 + pm1)) + c1_12 * sign * ((p2 - pm1) - T(3.0) * (p1 - p0));
}
template <typename T>
__device__ __forceinline__ T i2(T p0, T p1) { return T(0.5) * (p1 + p0); }
// d i5 / d operand d (0..5 = pm2 .. p3) at fixed sign
template <typename T>
__device__ __forceinline__ T dc5(int d, T s) {
    const T a = T(1.0 / 60.0);
    switch (d) {
    case 0: return a + a * s;
    case 1: return -T(2.0 / 15.0) - T(5.0) * a * s;
    case 2: return T(37.0 / 60.0) + T(10.0) * a * s;
    case 3: return T(37.0 / 60.0) - T(10.0) * a * s;
    case 4: return -T(2.0 / 15.0) + T(5.0) * a * s;
    default: return a - a * s;
    }
}
// d i3 / d operand d (0..3 = pm1 .. p2) at fixed sign
template <typename T>
__device__ __forceinline__ T dc3(int d, T s) {
    const T a = T(1.0 / 12.0);
    switch (d) {
    case 0: return -a - a * s;
    case 1: return T(7.0 / 12.0) + T(3.0) * a * s;
    case 2: return T(7.0 / 12.0) - T(3.0) * a * s;
    default: return -a + a * s;
    }
}

template <typename T>
struct AdvAdjArgs {
    const T *U, *V, *W, *TH;      // the stage's current fields (forward)
    const T *UB, *VB, *WB, *THB;  // basic state
    const T *PU, *PV, *PW, *PTH;  // adjoints of the forecast buffers
    T *CU, *CV, *CW, *CTH;        // adjoints of the current buffers (accumulated)
    int il, ih, jl, jh;
};

// Transpose of advect_uvw_velocity + advect_theta over the box at fixed upwind signs.  With A = step *
// (adjoint of the forecast) on the written points, a face flux F enters the tendencies of its two cells
// with opposite signs, so its adjoint is a difference of A; F = a * Phi_bar(sign a) + (abar + a) *
// Phi'(sign(abar + a)) then returns (abar + a) * dPhi'/d(operand) to the advected field and
// (Phi_bar + Phi') * da/d(velocity) to the advecting velocity.  One thread gathers every term that
// reaches its point.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_advect(Geo g, Cols c, AdvAdjArgs<T> a, T step) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    const int NZ = g.NZ;
    const int il = a.il, ih = a.ih, jl = a.jl, jh = a.jh;
    if (k >= NZ || j >= g.NY) return;
    if (i < il - 3 || i >= ih + 3 || j < jl - 3 || j >= jh + 3) return;
    auto X = [&](int ii, int jj, int kk) -> long long { return IX(ii, jj, kk); };
    auto inb = [&](int ii, int jj, int kk) { return ii >= il && ii < ih && jj >= jl && jj < jh && kk >= 1 && kk <= NZ - 2; };
    auto AU = [&](int ii, int jj, int kk) -> T { return inb(ii, jj, kk) ? step * a.PU[X(ii, jj, kk)] : T(0); };
    auto AV = [&](int ii, int jj, int kk) -> T { return inb(ii, jj, kk) ? step * a.PV[X(ii, jj, kk)] : T(0); };
    auto ATH = [&](int ii, int jj, int kk) -> T { return inb(ii, jj, kk) ? step * a.PTH[X(ii, jj, kk)] : T(0); };
    auto AW = [&](int ii, int jj, int kk) -> T { return inb(ii, jj, kk) && kk >= 2 ? step * a.PW[X(ii, jj, kk)] : T(0); };
    const T dtx = T(g.dtx), dty = T(g.dty), dt = T(g.dt);
    auto DTZ = [&](int kk) -> T { return T(g.dtz) * T(c.mu[kk]); };
    // adjoint of the top face kf of a field whose scaled tendency adjoint is A:
    // tend(k) -= DTZ(k) (rhow[k+1] top(k) - rhow[k] top(k-1)) one_d_rhou[k]
    auto zface = [&](auto A, int ii, int jj, int kf) -> T {
        return -DTZ(kf) * T(c.rhow[kf + 1]) * T(c.one_d_rhou[kf]) * A(ii, jj, kf) +
               DTZ(kf + 1) * T(c.rhow[kf + 1]) * T(c.one_d_rhou[kf + 1]) * A(ii, jj, kf + 1);
    };
    auto I5x = [&](const T *F, int e, int jj, int kk, T s) -> T {
        return i5(F[X(e - 2, jj, kk)], F[X(e - 1, jj, kk)], F[X(e, jj, kk)], F[X(e + 1, jj, kk)], F[X(e + 2, jj, kk)], F[X(e + 3, jj, kk)], s);
    };
    auto I5y = [&](const T *F, int ii, int f, int kk, T s) -> T {
        return i5(F[X(ii, f - 2, kk)], F[X(ii, f - 1, kk)], F[X(ii, f, kk)], F[X(ii, f + 1, kk)], F[X(ii, f + 2, kk)], F[X(ii, f + 3, kk)], s);
    };
    auto Iz = [&](const T *F, int ii, int jj, int kf, T s) -> T {
        if (kf == 1 || kf == NZ - 2) return i2(F[X(ii, jj, kf)], F[X(ii, jj, kf + 1)]);
        return i3(F[X(ii, jj, kf - 1)], F[X(ii, jj, kf)], F[X(ii, jj, kf + 1)], F[X(ii, jj, kf + 2)], s);
    };
    auto dIz = [&](int kf, int kk, T s) -> T {
        if (kf == 1 || kf == NZ - 2) return (kk == kf || kk == kf + 1) ? T(0.5) : T(0);
        const int d = kk - kf + 1;
        return (d >= 0 && d <= 3) ? dc3<T>(d, s) : T(0);
    };
    const bool krow = k >= 1 && k <= NZ - 2;
    const bool jrow = j >= jl && j < jh;
    const bool icol = i >= il && i < ih;
    T gu = T(0), gv = T(0), gw = T(0), gth = T(0);

    // ---- east faces e (between columns e and e+1), e in [il-1, ih-1], as advected-field stencils ----
    if (jrow && krow) {
        for (int e = max(i - 3, il - 1); e <= min(i + 2, ih - 1); e++) {
            const int d = i - e + 2;
            {  // u_east: ua = (U(e+1) + U(e)) / 2
                const T fh = dtx * (AU(e + 1, j, k) - AU(e, j, k));
                const T ua = T(0.5) * (a.U[X(e + 1, j, k)] + a.U[X(e, j, k)]), uba = T(0.5) * (a.UB[X(e + 1, j, k)] + a.UB[X(e, j, k)]);
                const T wt = uba + ua, s = sgn(wt);
                gu += fh * wt * dc5<T>(d, s);
                if (e == i - 1 || e == i) gu += T(0.5) * fh * (I5x(a.UB, e, j, k, sgn(ua)) + I5x(a.U, e, j, k, s));
            }
            {  // v_east: ua = (U(e+1, j) + U(e+1, j-1)) / 2
                const T fh = dtx * (AV(e + 1, j, k) - AV(e, j, k));
                const T ua = T(0.5) * (a.U[X(e + 1, j, k)] + a.U[X(e + 1, j - 1, k)]);
                const T uba = T(0.5) * (a.UB[X(e + 1, j, k)] + a.UB[X(e + 1, j - 1, k)]);
                const T wt = uba + ua;
                gv += fh * wt * dc5<T>(d, sgn(wt));
            }
            {  // s_east: U1 = U(e+1)
                const T fh = dtx * (ATH(e + 1, j, k) - ATH(e, j, k));
                const T wt = a.UB[X(e + 1, j, k)] + a.U[X(e + 1, j, k)];
                gth += fh * wt * dc5<T>(d, sgn(wt));
            }
        }
    }
    // ---- U as the advecting velocity of v (faces (i-1, j), (i-1, j+1)) and of theta (face i-1); Coriolis ----
    if (krow) {
        const int e = i - 1;
        if (e >= il - 1 && e <= ih - 1) {
            for (int jp = j; jp <= j + 1; jp++) {
                if (jp < jl || jp >= jh) continue;
                const T fh = dtx * (AV(e + 1, jp, k) - AV(e, jp, k));
                const T ua = T(0.5) * (a.U[X(e + 1, jp, k)] + a.U[X(e + 1, jp - 1, k)]);
                const T uba = T(0.5) * (a.UB[X(e + 1, jp, k)] + a.UB[X(e + 1, jp - 1, k)]);
                gu += T(0.5) * fh * (I5x(a.VB, e, jp, k, sgn(ua)) + I5x(a.V, e, jp, k, sgn(uba + ua)));
            }
            if (jrow) {
                const T fh = dtx * (ATH(i, j, k) - ATH(e, j, k));
                const T U1 = a.U[X(i, j, k)], UB1 = a.UB[X(i, j, k)];
                gu += fh * (I5x(a.THB, e, j, k, sgn(U1)) + I5x(a.TH, e, j, k, sgn(UB1 + U1)));
            }
        }
        // v: tend(i', j') -= dt f_j' (U(i',j') + U(i'+1,j') + U(i',j'-1) + U(i'+1,j'-1)) / 4
        gu -= T(0.25) * dt *
              (T(c.f[j]) * (AV(i, j, k) + AV(i - 1, j, k)) + (j + 1 < g.NY ? T(c.f[j + 1]) * (AV(i, j + 1, k) + AV(i - 1, j + 1, k)) : T(0)));
        // u: tend(i', j') += dt f_j' (V(i',j') + V(i',j'+1) + V(i'-1,j') + V(i'-1,j'+1)) / 4
        gv += T(0.25) * dt *
              (T(c.f[j]) * (AU(i, j, k) + AU(i + 1, j, k)) + (j >= 1 ? T(c.f[j - 1]) * (AU(i, j - 1, k) + AU(i + 1, j - 1, k)) : T(0)));
    }
    // ---- north faces f (between rows f and f+1), f in [jl-1, jh-1], as advected-field stencils ----
    if (icol && krow) {
        for (int f = max(j - 3, jl - 1); f <= min(j + 2, jh - 1); f++) {
            const int d = j - f + 2;
            {  // u_north: va = (V(i, f+1) + V(i-1, f+1)) / 2
                const T fh = dty * (AU(i, f + 1, k) - AU(i, f, k));
                const T va = T(0.5) * (a.V[X(i, f + 1, k)] + a.V[X(i - 1, f + 1, k)]);
                const T vba = T(0.5) * (a.VB[X(i, f + 1, k)] + a.VB[X(i - 1, f + 1, k)]);
                const T wt = vba + va;
                gu += fh * wt * dc5<T>(d, sgn(wt));
            }
            {  // v_north: va = (V(f+1) + V(f)) / 2
                const T fh = dty * (AV(i, f + 1, k) - AV(i, f, k));
                const T va = T(0.5) * (a.V[X(i, f + 1, k)] + a.V[X(i, f, k)]), vba = T(0.5) * (a.VB[X(i, f + 1, k)] + a.VB[X(i, f, k)]);
                const T wt = vba + va, s = sgn(wt);
                gv += fh * wt * dc5<T>(d, s);
                if (f == j - 1 || f == j) gv += T(0.5) * fh * (I5y(a.VB, i, f, k, sgn(va)) + I5y(a.V, i, f, k, s));
            }
            {  // s_north: V1 = V(i, f+1)
                const T fh = dty * (ATH(i, f + 1, k) - ATH(i, f, k));
                const T wt = a.VB[X(i, f + 1, k)] + a.V[X(i, f + 1, k)];
                gth += fh * wt * dc5<T>(d, sgn(wt));
            }
        }
    }
    // ---- V as the advecting velocity of u (north faces of columns i, i+1 at row j-1) and of theta ----
    if (krow) {
        const int f = j - 1;
        if (f >= jl - 1 && f <= jh - 1) {
            for (int ip = i; ip <= i + 1; ip++) {
                if (ip < il || ip >= ih) continue;
                const T fh = dty * (AU(ip, f + 1, k) - AU(ip, f, k));
                const T va = T(0.5) * (a.V[X(ip, f + 1, k)] + a.V[X(ip - 1, f + 1, k)]);
                const T vba = T(0.5) * (a.VB[X(ip, f + 1, k)] + a.VB[X(ip - 1, f + 1, k)]);
                gv += T(0.5) * fh * (I5y(a.UB, ip, f, k, sgn(va)) + I5y(a.U, ip, f, k, sgn(vba + va)));
            }
            if (icol) {
                const T fh = dty * (ATH(i, j, k) - ATH(i, f, k));
                const T V1 = a.V[X(i, j, k)], VB1 = a.VB[X(i, j, k)];
                gv += fh * (I5y(a.THB, i, f, k, sgn(V1)) + I5y(a.TH, i, f, k, sgn(VB1 + V1)));
            }
        }
    }
    // ---- top faces kf (between levels kf and kf+1), kf in [1, NZ-2], as advected-field stencils ----
    if (icol && jrow) {
        for (int kf = max(k - 2, 1); kf <= min(k + 1, NZ - 2); kf++) {
            {  // u_top: wa = (W(i, kf+1) + W(i-1, kf+1)) / 2
                const T fh = zface(AU, i, j, kf);
                const T wa = T(0.5) * (a.W[X(i, j, kf + 1)] + a.W[X(i - 1, j, kf + 1)]);
                const T wba = T(0.5) * (a.WB[X(i, j, kf + 1)] + a.WB[X(i - 1, j, kf + 1)]);
                const T wt = wba + wa;
                gu += fh * wt * dIz(kf, k, sgn(wt));
            }
            {  // v_top: wa = (W(j, kf+1) + W(j-1, kf+1)) / 2
                const T fh = zface(AV, i, j, kf);
                const T wa = T(0.5) * (a.W[X(i, j, kf + 1)] + a.W[X(i, j - 1, kf + 1)]);
                const T wba = T(0.5) * (a.WB[X(i, j, kf + 1)] + a.WB[X(i, j - 1, kf + 1)]);
                const T wt = wba + wa;
                gv += fh * wt * dIz(kf, k, sgn(wt));
            }
            {  // s_top: W1 = W(kf+1)
                const T fh = zface(ATH, i, j, kf);
                const T wt = a.WB[X(i, j, kf + 1)] + a.W[X(i, j, kf + 1)];
                gth += fh * wt * dIz(kf, k, sgn(wt));
            }
        }
    }
    // ---- W as the advecting velocity (face k-1) of u (columns i, i+1), v (rows j, j+1) and theta ----
    {
        const int kf = k - 1;
        if (kf >= 1 && kf <= NZ - 2) {
            if (jrow) {
                for (int ip = i; ip <= i + 1; ip++) {
                    if (ip < il || ip >= ih) continue;
                    const T fh = zface(AU, ip, j, kf);
                    const T wa = T(0.5) * (a.W[X(ip, j, k)] + a.W[X(ip - 1, j, k)]), wba = T(0.5) * (a.WB[X(ip, j, k)] + a.WB[X(ip - 1, j, k)]);
                    gw += T(0.5) * fh * (Iz(a.UB, ip, j, kf, sgn(wa)) + Iz(a.U, ip, j, kf, sgn(wba + wa)));
                }
            }
            if (icol) {
                for (int jp = j; jp <= j + 1; jp++) {
                    if (jp < jl || jp >= jh) continue;
                    const T fh = zface(AV, i, jp, kf);
                    const T wa = T(0.5) * (a.W[X(i, jp, k)] + a.W[X(i, jp - 1, k)]), wba = T(0.5) * (a.WB[X(i, jp, k)] + a.WB[X(i, jp - 1, k)]);
                    gw += T(0.5) * fh * (Iz(a.VB, i, jp, kf, sgn(wa)) + Iz(a.V, i, jp, kf, sgn(wba + wa)));
                }
            }
            if (icol && jrow) {
                const T fh = zface(ATH, i, j, kf);
                const T W1 = a.W[X(i, j, k)], WB1 = a.WB[X(i, j, k)];
                gw += fh * (Iz(a.THB, i, j, kf, sgn(W1)) + Iz(a.TH, i, j, kf, sgn(WB1 + W1)));
            }
        }
        if (icol && jrow && k >= 1) {
            // Z_ADVECT_BASE (advection.cpp:16-18): W(k) enters the base-state term of levels k-1 and k
            const T coef = T(0.5) * T(c.rhow[k]) * (T(c.tb[k]) - T(c.tb[k - 1]));
            gw -= coef * (DTZ(k - 1) * T(c.one_d_rhou[k - 1]) * ATH(i, j, k - 1) + DTZ(k) * T(c.one_d_rhou[k]) * ATH(i, j, k));
        }
        // buoyancy (advection.cpp:329-337): WP(k') = WM + step dt g (th(k')/tbv(k') + th(k'-1)/tbv(k'-1)) / 2
        if (icol && jrow) gth += dt * T(kGrav) * T(0.5) * (AW(i, j, k) + (k + 1 < NZ ? AW(i, j, k + 1) : T(0))) / T(c.tbv[k]);
    }
    const long long p = X(i, j, k);
    a.CU[p] += gu;
    a.CV[p] += gv;
    a.CW[p] += gw;
    a.CTH[p] += gth;
}

// forecast = M + step * tendency on the box: the M adjoint takes the forecast adjoint, which the overwrite
// then clears
template <typename T>
__global__ void k_adj_advect_finish(Geo g, T *MU, T *MV, T *MW, T *MTH, T *PU, T *PV, T *PW, T *PTH, int il, int jl, int jh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + jl;
    const int i = blockIdx.z + il;
    if (k > g.NZ - 2 || j >= jh) return;
    const long long p = IX(i, j, k);
    MU[p] += PU[p];
    PU[p] = T(0);
    MV[p] += PV[p];
    PV[p] = T(0);
    MTH[p] += PTH[p];
    PTH[p] = T(0);
    if (k >= 2) {
        MW[p] += PW[p];
        PW[p] = T(0);
    }
}

// transpose of the pressure-gradient force (solver.cpp:123-135) over [1,NX-1) x [1,NY-1): the forecast
// adjoints pass through, PI's adjoint gathers the gradient stencil
template <typename T>
__global__ void k_adj_pgf(Geo g, Cols c, const T *__restrict__ UH, const T *__restrict__ VH, const T *__restrict__ WH,
                          T *__restrict__ PH, T step) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.NZ || j >= g.NY) return;
    auto inr = [&](int ii, int jj, int kk) { return ii >= 1 && ii < g.NX - 1 && jj >= 1 && jj < g.NY - 1 && kk >= 1 && kk <= g.NZ - 2; };
    const long long p = IX(i, j, k);
    const T sx = step * T(g.dtx), sy = step * T(g.dty);
    T acc = T(0);
    if (inr(i, j, k)) acc -= sx * UH[p] + sy * VH[p];
    if (inr(i + 1, j, k)) acc += sx * UH[p + g.NYNZ];
    if (inr(i, j + 1, k)) acc += sy * VH[p + g.NZ];
    if (inr(i, j, k) && k >= 2) acc -= step * (T(g.dtz) * T(c.mw[k])) * WH[p];
    if (inr(i, j, k + 1) && k + 1 >= 2) acc += step * (T(g.dtz) * T(c.mw[k + 1])) * WH[p + 1];
    PH[p] += acc;
}

struct AdjList {
    void *f[5];
    int n;
};

// transpose of upper_lower_boundaries (from i = 1, j = 1): level 0 and NZ-1 adjoints move onto 1 and NZ-2
// (nfill fields), w's lid levels are cleared
template <typename T>
__global__ void k_adj_lid(Geo g, AdjList fl, int nfill) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y + 1;
    if (j >= g.NY) return;
    const long long p = IX(i, j, 0);
    for (int f = 0; f < fl.n; f++) {
        T *A = (T *)fl.f[f];
        if (f < nfill) {
            A[p + 1] += A[p];
            A[p] = T(0);
            A[p + g.NZ - 2] += A[p + g.NZ - 1];
            A[p + g.NZ - 1] = T(0);
        } else {
            A[p] = T(0);
            A[p + 1] = T(0);
            A[p + g.NZ - 1] = T(0);
        }
    }
}

// transpose of periodic_boundaries_east_west: halo adjoints return to their source columns
template <typename T>
__global__ void k_adj_periodic(Geo g, AdjList fl) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.NYNZ) return;
    const long long s = g.NYNZ;
    for (int f = 0; f < fl.n; f++) {
        T *A = (T *)fl.f[f];
        for (int h = 0; h < 3; h++) {
            A[q + (long long)(g.NX - 6 + h) * s] += A[q + (long long)h * s];
            A[q + (long long)h * s] = T(0);
            A[q + (long long)(3 + h) * s] += A[q + (long long)(g.NX - 3 + h) * s];
            A[q + (long long)(g.NX - 3 + h) * s] = T(0);
        }
    }
}

// sponge rows (a pointwise scaling, its own transpose): rows [0,js) and [NY-jn,NY), i in [1,NX), k in [1,NZ)
template <typename T>
__global__ void k_adj_sponge(Geo g, AdjList fl, const double *__restrict__ cs, const double *__restrict__ cn, int js, int jn) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int r = blockIdx.y, i = blockIdx.z + 1;
    if (k >= g.NZ) return;
    int j;
    T coef;
    if (r < js) {
        j = r;
        coef = T(cs[r]);
    } else {
        j = g.NY - jn + (r - js);
        coef = T(cn[r - js]);
    }
    const long long p = IX(i, j, k);
    for (int f = 0; f < fl.n; f++) {
        T *A = (T *)fl.f[f];
        A[p] -= coef * A[p];
    }
}

// Rayleigh damping (pointwise scaling, its own transpose) over [1,NX-1) x [1,NY-1) x [level, NZ)
template <typename T>
__global__ void k_adj_rayleigh(Geo g, const double *__restrict__ coef, T *U, T *V, T *W, T *TH, int level) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + level;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k >= g.NZ || j >= g.NY - 1) return;
    const long long p = IX(i, j, k);
    const T cu = T(coef[k]), cw = T(coef[g.NZ + k]);
    U[p] -= cu * U[p];
    V[p] -= cu * V[p];
    W[p] -= cw * W[p];
    TH[p] -= cu * TH[p];
}

template <typename T>
__global__ void k_add_clear(T *__restrict__ dst, T *__restrict__ src, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        dst[p] += src[p];
        src[p] = T(0);
    }
}
template <typename T>
__global__ void k_sum3(double *__restrict__ out, const T *__restrict__ a, const T *__restrict__ b, const T *__restrict__ c, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = double(a[p]) + double(b[p]) + double(c[p]);
}
template <typename T>
__global__ void k_from_double(T *__restrict__ out, const double *__restrict__ in, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = in ? T(in[p]) : T(0);
}
template <typename T>
__global__ void k_zero(T *__restrict__ a, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) a[p] = T(0);
}

}  // namespace

// Trajectory storage and adjoint buffers of one context (allocated by dyn_adjoint_record).
struct AdjState {
    int nsteps = 0, every = 0;
    size_t field_bytes = 0;
    std::vector<void *> snap;  // snapshot s, field f, role r at snap[(s * 4 + f) * 3 + r]; last = the final state
    std::vector<void *> rec;   // segment step n, stage st, field f at rec[(n * 3 + st) * 4 + f]
    void *adj[kAdjFields][3] = {{nullptr}};
    void *pihat = nullptr, *r1 = nullptr, *r2 = nullptr;
    double *stage_dbl = nullptr;  // host-transfer staging (double)
    int rec_base = 0;             // first recorded step of the segment in the record
};

namespace {

void adj_free(Ctx *c) {
    AdjState *s = c->adj;
    if (!s) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : s->snap) (void)hipFree(p);
    for (void *p : s->rec) (void)hipFree(p);
    for (auto &f : s->adj)
        for (void *p : f)
            if (p) (void)hipFree(p);
    void *ptrs[] = {s->pihat, s->r1, s->r2, s->stage_dbl};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete s;
    c->adj = nullptr;
}

int alloc_fields(std::vector<void *> &v, size_t count, size_t bytes) {
    v.assign(count, nullptr);
    for (auto &p : v) DYN_HIP(hipMalloc(&p, bytes));
    return DYN_OK;
}

int copy_state(Ctx *c, void *const *dst_or_null, void *const *src_or_null, bool to_snapshot) {
    AdjState *s = c->adj;
    for (int f = 0; f < kAdjFields; f++)
        for (int r = 0; r < 3; r++) {
            void *prog = c->prog[kAdjIds[f]][r];
            void *snap = to_snapshot ? dst_or_null[f * 3 + r] : src_or_null[f * 3 + r];
            if (to_snapshot) DYN_HIP(hipMemcpyAsync(snap, prog, s->field_bytes, hipMemcpyDeviceToDevice, c->stream));
            else DYN_HIP(hipMemcpyAsync(prog, snap, s->field_bytes, hipMemcpyDeviceToDevice, c->stream));
        }
    return DYN_OK;
}

// stage hook of the recomputation: the current buffers at the start of every stage
int record_hook(Ctx *c, int stage, void *arg) {
    const int n = *(const int *)arg;
    AdjState *s = c->adj;
    for (int f = 0; f < kAdjFields; f++)
        DYN_HIP(hipMemcpyAsync(s->rec[((size_t)n * 3 + stage) * 4 + f], c->prog[kAdjIds[f]][DYN_ROLE_CUR], s->field_bytes,
                               hipMemcpyDeviceToDevice, c->stream));
    return DYN_OK;
}

template <typename T>
int bc_adjoint(Ctx *c) {
    AdjState *s = c->adj;
    const Geo &g = c->geo;
    void *U = s->adj[0][DYN_ROLE_P], *V = s->adj[1][DYN_ROLE_P], *W = s->adj[2][DYN_ROLE_P], *TH = s->adj[3][DYN_ROLE_P];
    Timed _t(c, "adj_bc", (double)c->npts);
    hipLaunchKernelGGL(k_adj_lid<T>, dim3((g.NY - 1 + 63) / 64, g.NX - 1), dim3(64), 0, c->stream, g, AdjList{{U, V, TH, W}, 4}, 3);
    DYN_LAUNCH_CHECK();
    const long long n = g.NYNZ;
    hipLaunchKernelGGL(k_adj_periodic<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, g,
                       AdjList{{U, V, s->pihat, TH, W}, 5});
    DYN_LAUNCH_CHECK();
    const int rows = c->g.sponge_south + c->g.sponge_north;
    if (rows > 0) {
        hipLaunchKernelGGL(k_adj_sponge<T>, dim3((g.NZ - 1 + 63) / 64, rows, g.NX - 1), dim3(64), 0, c->stream, g, AdjList{{U, V, TH, W}, 4},
                           c->sponge_s, c->sponge_n, c->g.sponge_south, c->g.sponge_north);
        DYN_LAUNCH_CHECK();
    }
    return DYN_OK;
}

template <typename T>
int stage_adjoint(Ctx *c, int n, int stage, double step) {
    AdjState *s = c->adj;
    const Geo &g = c->geo;
    const int NX = c->g.nx, NY = c->g.ny;
    auto P = [&](int f) { return (T *)s->adj[f][DYN_ROLE_P]; };
    int rc = bc_adjoint<T>(c);
    if (rc) return rc;
    {
        Timed _t(c, "adj_pgf", (double)c->npts);
        hipLaunchKernelGGL(k_adj_pgf<T>, dim3((g.NZ + 63) / 64, (g.NY + 3) / 4, g.NX), dim3(64, 4), 0, c->stream, g, c->cols,
                           (const T *)P(0), (const T *)P(1), (const T *)P(2), (T *)s->pihat, T(step));
        DYN_LAUNCH_CHECK();
    }
    rc = launch_pressure_adjoint(c, step, s->pihat, P(0), P(1), P(2), s->r1, s->r2);
    if (rc) return rc;
    {  // periodic_uvw_boundaries of solve_pressure
        const long long nq = g.NYNZ;
        hipLaunchKernelGGL(k_adj_periodic<T>, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, c->stream, g,
                           AdjList{{P(0), P(1), P(2)}, 3});
        DYN_LAUNCH_CHECK();
    }
    AdvAdjArgs<T> a{};
    const size_t rb = ((size_t)(n - s->rec_base) * 3 + stage) * 4;
    a.U = (const T *)s->rec[rb + 0];
    a.V = (const T *)s->rec[rb + 1];
    a.W = (const T *)s->rec[rb + 2];
    a.TH = (const T *)s->rec[rb + 3];
    a.UB = (const T *)c->bar[0];
    a.VB = (const T *)c->bar[1];
    a.WB = (const T *)c->bar[2];
    a.THB = (const T *)c->bar[3];
    a.PU = P(0);
    a.PV = P(1);
    a.PW = P(2);
    a.PTH = P(3);
    a.CU = (T *)s->adj[0][DYN_ROLE_CUR];
    a.CV = (T *)s->adj[1][DYN_ROLE_CUR];
    a.CW = (T *)s->adj[2][DYN_ROLE_CUR];
    a.CTH = (T *)s->adj[3][DYN_ROLE_CUR];
    a.il = 3;
    a.ih = NX - 3;
    a.jl = 3;
    a.jh = NY - 3;
    {
        Timed _t(c, "adj_advect", (double)c->npts);
        hipLaunchKernelGGL(k_adj_advect<T>, dim3((g.NZ + 63) / 64, (g.NY + 3) / 4, g.NX), dim3(64, 4), 0, c->stream, g, c->cols, a, T(step));
        DYN_LAUNCH_CHECK();
        auto M = [&](int f) { return (T *)s->adj[f][DYN_ROLE_M]; };
        hipLaunchKernelGGL(k_adj_advect_finish<T>, dim3((g.NZ - 2 + 63) / 64, (a.jh - a.jl + 3) / 4, a.ih - a.il), dim3(64, 4), 0, c->stream, g,
                           M(0), M(1), M(2), M(3), P(0), P(1), P(2), P(3), a.il, a.jl, a.jh);
        DYN_LAUNCH_CHECK();
    }
    return DYN_OK;
}

void swap_cur_p(AdjState *s) {
    for (int f = 0; f < kAdjFields; f++) std::swap(s->adj[f][DYN_ROLE_CUR], s->adj[f][DYN_ROLE_P]);
}

// reverse of one forward step n (records of the step's stages in place)
template <typename T>
int step_adjoint(Ctx *c, int n) {
    AdjState *s = c->adj;
    const Geo &g = c->geo;
    const long long npts = (long long)c->npts;
    const dim3 blocks((unsigned)((npts + 255) / 256));
    // advance_outer: memcpy(M <- P) then swap(current, forecast)
    swap_cur_p(s);
    for (int f = 0; f < kAdjFields; f++) {
        hipLaunchKernelGGL(k_add_clear<T>, blocks, dim3(256), 0, c->stream, (T *)s->adj[f][DYN_ROLE_P], (T *)s->adj[f][DYN_ROLE_M], npts);
        DYN_LAUNCH_CHECK();
    }
    if (c->g.rayleigh_level < c->g.nz) {
        const int lev = c->g.rayleigh_level;
        hipLaunchKernelGGL(k_adj_rayleigh<T>, dim3((g.NZ - lev + 63) / 64, (g.NY - 2 + 3) / 4, g.NX - 2), dim3(64, 4), 0, c->stream, g,
                           c->raycoef, (T *)s->adj[0][DYN_ROLE_P], (T *)s->adj[1][DYN_ROLE_P], (T *)s->adj[2][DYN_ROLE_P],
                           (T *)s->adj[3][DYN_ROLE_P], lev);
        DYN_LAUNCH_CHECK();
    }
    int rc = bc_adjoint<T>(c);
    if (rc) return rc;
    const double steps[] = {1. / 3., 0.5, 1.0};
    for (int st = 2; st >= 0; st--) {
        if (st < 2) swap_cur_p(s);  // advance_inner
        rc = stage_adjoint<T>(c, n, st, steps[st]);
        if (rc) return rc;
    }
    return DYN_OK;
}

template <typename T>
int backward_t(Ctx *c, const double *const *seed, double *const *grad) {
    AdjState *s = c->adj;
    const long long npts = (long long)c->npts;
    const dim3 blocks((unsigned)((npts + 255) / 256));
    const size_t dbytes = sizeof(double) * (size_t)npts;
    for (int f = 0; f < kAdjFields; f++)
        for (int r = 0; r < 3; r++) {
            hipLaunchKernelGGL(k_zero<T>, blocks, dim3(256), 0, c->stream, (T *)s->adj[f][r], npts);
            DYN_LAUNCH_CHECK();
        }
    hipLaunchKernelGGL(k_zero<T>, blocks, dim3(256), 0, c->stream, (T *)s->pihat, npts);
    DYN_LAUNCH_CHECK();
    // seed: the adjoint of the final current buffers
    for (int f = 0; f < kAdjFields; f++) {
        if (!seed[f]) continue;
        DYN_HIP(hipMemcpyAsync(s->stage_dbl, seed[f], dbytes, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_from_double<T>, blocks, dim3(256), 0, c->stream, (T *)s->adj[f][DYN_ROLE_CUR], s->stage_dbl, npts);
        DYN_LAUNCH_CHECK();
    }
    const int nseg = (s->nsteps + s->every - 1) / s->every;
    for (int seg = nseg - 1; seg >= 0; seg--) {
        const int n0 = seg * s->every, n1 = std::min(s->nsteps, n0 + s->every);
        int rc = copy_state(c, nullptr, &s->snap[(size_t)seg * 12], false);
        if (rc) return rc;
        s->rec_base = n0;
        for (int n = n0; n < n1; n++) {
            int local = n - n0;
            rc = step_with_hook(c, record_hook, &local);
            if (rc) return rc;
        }
        for (int n = n1 - 1; n >= n0; n--) {
            rc = step_adjoint<T>(c, n);
            if (rc) return rc;
        }
    }
    // the initial fields are set into all three buffers: their gradient is the sum of the three adjoints
    for (int f = 0; f < kAdjFields; f++) {
        if (!grad[f]) continue;
        hipLaunchKernelGGL(k_sum3<T>, blocks, dim3(256), 0, c->stream, s->stage_dbl, (const T *)s->adj[f][0], (const T *)s->adj[f][1],
                           (const T *)s->adj[f][2], npts);
        DYN_LAUNCH_CHECK();
        DYN_HIP(hipMemcpyAsync(grad[f], s->stage_dbl, dbytes, hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(hipStreamSynchronize(c->stream));
    }
    // leave the context at the end of the forward run
    int rc = copy_state(c, nullptr, &s->snap[(size_t)nseg * 12], false);
    if (rc) return rc;
    DYN_HIP(hipStreamSynchronize(c->stream));
    return DYN_OK;
}

}  // namespace

void adjoint_release(Ctx *c) { adj_free(c); }

}  // namespace dyn

using namespace dyn;

extern "C" {

int dyn_adjoint_record(dyn_ctx *ctx, int nsteps, int every) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) {
        set_error("null context");
        return DYN_E_ARG;
    }
    if (c->g.microphysics != 0) {
        set_error("dyn_adjoint_record: the adjoint covers the dry step (microphysics option 0)");
        return DYN_E_UNSUPPORTED;
    }
    if (nsteps <= 0 || every <= 0) {
        set_error("dyn_adjoint_record: nsteps and every must be positive");
        return DYN_E_ARG;
    }
    adj_free(c);
    AdjState *s = new AdjState();
    c->adj = s;
    s->nsteps = nsteps;
    s->every = std::min(every, nsteps);
    s->field_bytes = c->elt * c->npts;
    const int nseg = (nsteps + s->every - 1) / s->every;
    int rc = alloc_fields(s->snap, (size_t)(nseg + 1) * 12, s->field_bytes);
    if (!rc) rc = alloc_fields(s->rec, (size_t)s->every * 12, s->field_bytes);
    if (rc) {
        adj_free(c);
        return rc;
    }
    for (auto &f : s->adj)
        for (auto &p : f) DYN_HIP(hipMalloc(&p, s->field_bytes));
    DYN_HIP(hipMalloc(&s->pihat, s->field_bytes));
    DYN_HIP(hipMalloc(&s->r1, s->field_bytes));
    DYN_HIP(hipMalloc(&s->r2, s->field_bytes));
    DYN_HIP(hipMalloc(&s->stage_dbl, sizeof(double) * c->npts));
    // forward pass with a snapshot every `every` steps and one of the final state
    for (int n = 0; n < nsteps; n++) {
        if (n % s->every == 0) {
            rc = copy_state(c, &s->snap[(size_t)(n / s->every) * 12], nullptr, true);
            if (rc) return rc;
        }
        rc = step_with_hook(c, nullptr, nullptr);
        if (rc) return rc;
    }
    rc = copy_state(c, &s->snap[(size_t)nseg * 12], nullptr, true);
    if (rc) return rc;
    DYN_HIP(hipStreamSynchronize(c->stream));
    return DYN_OK;
}

int dyn_adjoint_backward(dyn_ctx *ctx, const double *seed_u, const double *seed_v, const double *seed_w, const double *seed_th,
                         double *grad_u, double *grad_v, double *grad_w, double *grad_th) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) {
        set_error("null context");
        return DYN_E_ARG;
    }
    if (!c->adj) {
        set_error("dyn_adjoint_backward: no recorded trajectory (call dyn_adjoint_record first)");
        return DYN_E_STATE;
    }
    if (c->g.rayleigh_level < c->g.nz && c->ray_level_cached != c->g.rayleigh_level) {
        set_error("dyn_adjoint_backward: Rayleigh coefficients missing");
        return DYN_E_STATE;
    }
    const double *seed[kAdjFields] = {seed_u, seed_v, seed_w, seed_th};
    double *grad[kAdjFields] = {grad_u, grad_v, grad_w, grad_th};
    return c->prec == DYN_FP32 ? backward_t<float>(c, seed, grad) : backward_t<double>(c, seed, grad);
}

int dyn_adjoint_release(dyn_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DYN_E_ARG;
    adj_free(c);
    return DYN_OK;
}

}  // extern "C"
