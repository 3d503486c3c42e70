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

// element offsets are 32-bit: dyn_adjoint_record accepts fields of fewer than 2^31 points
#define IX(i, j, k) ((unsigned)(i) * (unsigned)g.NYNZ + (unsigned)(j) * (unsigned)g.NZ + (unsigned)(k))

constexpr int kAdjFields = 4;  // U V W TH
const int kAdjIds[kAdjFields] = {DYN_U, DYN_V, DYN_W, DYN_TH};

template <typename T>
__device__ __forceinline__ T sgn(T x) { return x > T(0) ? T(1) : T(-1); }
template <typename T>
__device__ __forceinline__ T i5(T pm2, T pm1, T p0, T p1, T p2, T p3, T sign) { return face_i5(pm2, pm1, p0, p1, p2, p3, sign); }
template <typename T>
__device__ __forceinline__ T i3(T pm1, T p0, T p1, T p2, T sign) { return face_i3(pm1, p0, p1, p2, sign); }
template <typename T>
__device__ __forceinline__ T i2(T p0, T p1) { return T(0.5) * (p1 + p0); }
// thread -> point (i, j, k) of the whole array, k fastest (every lane busy whatever NZ is)
__device__ __forceinline__ bool point_of(const Geo &g, int &i, int &j, int &k) {
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)g.NX * (unsigned)g.NYNZ) return false;
    i = (int)(q / (unsigned)g.NYNZ);
    const int r = (int)(q - (unsigned)i * (unsigned)g.NYNZ);
    j = r / g.NZ;
    k = r - j * g.NZ;
    return true;
}
inline dim3 point_grid(const Geo &g) { return dim3((unsigned)(((long long)g.NX * g.NYNZ + 255) / 256)); }

template <typename T>
struct AdvAdjArgs {
    const T *U, *V, *W, *TH;      // the stage's current fields (forward)
    const T *UB, *VB, *WB, *THB;  // basic state
    const T *PU, *PV, *PW, *PTH;  // adjoints of the forecast buffers
    T *CU, *CV, *CW, *CTH;        // adjoints of the current buffers (accumulated)
    int il, ih, jl, jh;
};

// Transpose of advect_uvw_velocity + advect_theta over the box at fixed upwind signs, in two passes.
// With A = step * (adjoint of the forecast) on the written points, a face flux F enters the tendencies of
// its two cells with opposite signs, so its adjoint fh is a difference of A.  F = a Phi_bar(sign a) +
// (abar + a) Phi'(sign(abar + a)) returns (abar + a) dPhi'/d(operand) to the advected field and
// (Phi_bar + Phi') da/d(velocity) to the advecting velocity.  Pass 1 (k_adj_faces) evaluates, once per
// face, for the nine face kinds (u, v, theta on east, north and top faces):
//   Aw = fh (abar + a), s = sign(abar + a) (one byte), G = fh (Phi_bar + Phi');
// pass 2 (k_adj_gather) gathers every term that reaches a point, the interpolant's derivative being
// c_d + s u_d (centred and upwind coefficients of operand d).
enum { FEU = 0, FEV, FET, FNU, FNV, FNT, FZU, FZV, FZT, NFACE };
struct AdjFaces {
    void *f[2 * NFACE];         // [kind][Aw, G], each a full field indexed by the face's low cell
    signed char *sgn[NFACE];    // [kind] upwind sign of the face
};

template <typename T>
struct AdjCommon {
    const Geo &g;
    const Cols &c;
    const AdvAdjArgs<T> &a;
    T step;
    __device__ long long X(int ii, int jj, int kk) const { return IX(ii, jj, kk); }
    __device__ bool inb(int ii, int jj, int kk) const {
        return ii >= a.il && ii < a.ih && jj >= a.jl && jj < a.jh && kk >= 1 && kk <= g.NZ - 2;
    }
    __device__ T AU(int ii, int jj, int kk) const { return inb(ii, jj, kk) ? step * a.PU[X(ii, jj, kk)] : T(0); }
    __device__ T AV(int ii, int jj, int kk) const { return inb(ii, jj, kk) ? step * a.PV[X(ii, jj, kk)] : T(0); }
    __device__ T ATH(int ii, int jj, int kk) const { return inb(ii, jj, kk) ? step * a.PTH[X(ii, jj, kk)] : T(0); }
    __device__ T AW(int ii, int jj, int kk) const { return inb(ii, jj, kk) && kk >= 2 ? step * a.PW[X(ii, jj, kk)] : T(0); }
    // the same values from an unconditional load at the index clamped into the array, then a select
    __device__ long long XC(int ii, int jj, int kk) const {
        return IX(min(max(ii, 0), g.NX - 1), min(max(jj, 0), g.NY - 1), min(max(kk, 0), g.NZ - 1));
    }
    __device__ T AUc(int ii, int jj, int kk) const {
        const T v = step * a.PU[XC(ii, jj, kk)];
        return inb(ii, jj, kk) ? v : T(0);
    }
    __device__ T AVc(int ii, int jj, int kk) const {
        const T v = step * a.PV[XC(ii, jj, kk)];
        return inb(ii, jj, kk) ? v : T(0);
    }
    __device__ T AWc(int ii, int jj, int kk) const {
        const T v = step * a.PW[XC(ii, jj, kk)];
        return inb(ii, jj, kk) && kk >= 2 ? v : T(0);
    }
    __device__ T ATHc(int ii, int jj, int kk) const {
        const T v = step * a.PTH[XC(ii, jj, kk)];
        return inb(ii, jj, kk) ? v : T(0);
    }
    __device__ T DTZ(int kk) const { return T(g.dtz) * T(c.mu[kk]); }
    // tend(k) -= DTZ(k) (rhow[k+1] top(k) - rhow[k] top(k-1)) one_d_rhou[k]: adjoint of the top face kf
    __device__ T zcoef_lo(int kf) const { return -DTZ(kf) * T(c.rhow[kf + 1]) * T(c.one_d_rhou[kf]); }
    __device__ T zcoef_hi(int kf) const { return DTZ(kf + 1) * T(c.rhow[kf + 1]) * T(c.one_d_rhou[kf + 1]); }
};

template <typename T>
__global__ void __launch_bounds__(256) k_adj_gather(Geo g, Cols c, AdvAdjArgs<T> a, AdjFaces fa, T step) {
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    const int NZ = g.NZ, il = a.il, ih = a.ih, jl = a.jl, jh = a.jh;
    if (i < il - 3 || i >= ih + 3 || j < jl - 3 || j >= jh + 3) return;
    const AdjCommon<T> m{g, c, a, step};
    // q = 0: Aw; q = 2: G; S: the face's upwind sign
    auto F = [&](int kind, int q, int ii, int jj, int kk) -> T { return ((const T *)fa.f[2 * kind + q / 2])[IX(ii, jj, kk)]; };
    auto S = [&](int kind, int ii, int jj, int kk) -> T { return T(fa.sgn[kind][IX(ii, jj, kk)]); };
    // Aw (c + s u) of a face for operand coefficients c, u
    auto stencil = [&](int kind, int ii, int jj, int kk, T cc, T uu) -> T { return F(kind, 0, ii, jj, kk) * (cc + S(kind, ii, jj, kk) * uu); };
    // i5 = sum_d (c6[d] + sign * u5[d]) p_d and i3 = sum_d (c4[d] + sign * u3[d]) p_d
    const T c6[6] = {T(1.0 / 60.0), -T(2.0 / 15.0), T(37.0 / 60.0), T(37.0 / 60.0), -T(2.0 / 15.0), T(1.0 / 60.0)};
    const T u5[6] = {T(1.0 / 60.0), -T(5.0 / 60.0), T(10.0 / 60.0), -T(10.0 / 60.0), T(5.0 / 60.0), -T(1.0 / 60.0)};
    const T c4[4] = {-T(1.0 / 12.0), T(7.0 / 12.0), T(7.0 / 12.0), -T(1.0 / 12.0)};
    const T u3[4] = {-T(1.0 / 12.0), T(3.0 / 12.0), -T(3.0 / 12.0), T(1.0 / 12.0)};
    const T dt = T(g.dt);
    const bool krow = k >= 1 && k <= NZ - 2, jrow = j >= jl && j < jh, icol = i >= il && i < ih;
    T gu = T(0), gv = T(0), gw = T(0), gth = T(0);
    // The three face loops run their six faces unrolled: every load of a loop is issued (clamped
    // indices) before any is used, and a face outside the loop's range contributes nothing (a select, so
    // whatever the face arrays hold there never enters).  A scheduling barrier after each loop keeps its
    // loads from being hoisted into the one before, which would raise the register count.
    const int jc = min(max(j, 0), g.NY - 1), kc = min(max(k, 1), NZ - 2);
    // east faces: stencils of u, v, theta; u as its own advecting velocity
    if (jrow && krow) {
        T aw[6][3], sg[6][3], ge[2];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const long long o = IX(min(max(i - 3 + q, 0), g.NX - 1), jc, kc);
            aw[q][0] = ((const T *)fa.f[2 * FEU])[o];
            aw[q][1] = ((const T *)fa.f[2 * FEV])[o];
            aw[q][2] = ((const T *)fa.f[2 * FET])[o];
            sg[q][0] = T(fa.sgn[FEU][o]);
            sg[q][1] = T(fa.sgn[FEV][o]);
            sg[q][2] = T(fa.sgn[FET][o]);
        }
        ge[0] = F(FEU, 2, max(i - 1, 0), j, k);
        ge[1] = F(FEU, 2, i, j, k);
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int e = i - 3 + q, d = 5 - q;  // d = i - e + 2
            if (e >= il - 1 && e <= ih - 1) {
                gu += aw[q][0] * (c6[d] + sg[q][0] * u5[d]);
                gv += aw[q][1] * (c6[d] + sg[q][1] * u5[d]);
                gth += aw[q][2] * (c6[d] + sg[q][2] * u5[d]);
                if (q == 2 || q == 3) gu += T(0.5) * ge[q - 2];
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (krow) {
        const int e = i - 1, ec = max(e, 0), jp = min(j + 1, g.NY - 1), jm = max(j - 1, 0);
        const T g0 = F(FEV, 2, ec, j, k), g1 = F(FEV, 2, ec, jp, k), g2 = F(FET, 2, ec, j, k);
        const T av0 = m.AVc(i, j, k), av1 = m.AVc(i - 1, j, k), av2 = m.AVc(i, j + 1, k), av3 = m.AVc(i - 1, j + 1, k);
        const T au0 = m.AUc(i, j, k), au1 = m.AUc(i + 1, j, k), au2 = m.AUc(i, j - 1, k), au3 = m.AUc(i + 1, j - 1, k);
        const T f0 = T(c.f[j]), fp = T(c.f[jp]), fm = T(c.f[jm]);
        if (e >= il - 1 && e <= ih - 1) {  // U(i, j) advects v at (e, j) and (e, j+1) and theta at e
            if (j >= jl && j < jh) gu += T(0.5) * g0;
            if (j + 1 >= jl && j + 1 < jh) gu += T(0.5) * g1;
            if (jrow) gu += g2;
        }
        // Coriolis: v tend -= dt f (U(i',j') + U(i'+1,j') + U(i',j'-1) + U(i'+1,j'-1)) / 4, u tend += dt f (V ...) / 4
        gu -= T(0.25) * dt * (f0 * (av0 + av1) + (j + 1 < g.NY ? fp * (av2 + av3) : T(0)));
        gv += T(0.25) * dt * (f0 * (au0 + au1) + (j >= 1 ? fm * (au2 + au3) : T(0)));
    }
    __builtin_amdgcn_sched_barrier(0);
    // north faces
    if (icol && krow) {
        T aw[6][3], sg[6][3], gn[2];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const long long o = IX(i, min(max(j - 3 + q, 0), g.NY - 1), kc);
            aw[q][0] = ((const T *)fa.f[2 * FNU])[o];
            aw[q][1] = ((const T *)fa.f[2 * FNV])[o];
            aw[q][2] = ((const T *)fa.f[2 * FNT])[o];
            sg[q][0] = T(fa.sgn[FNU][o]);
            sg[q][1] = T(fa.sgn[FNV][o]);
            sg[q][2] = T(fa.sgn[FNT][o]);
        }
        gn[0] = F(FNV, 2, i, max(j - 1, 0), k);
        gn[1] = F(FNV, 2, i, j, k);
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int f = j - 3 + q, d = 5 - q;  // d = j - f + 2
            if (f >= jl - 1 && f <= jh - 1) {
                gu += aw[q][0] * (c6[d] + sg[q][0] * u5[d]);
                gv += aw[q][1] * (c6[d] + sg[q][1] * u5[d]);
                gth += aw[q][2] * (c6[d] + sg[q][2] * u5[d]);
                if (q == 2 || q == 3) gv += T(0.5) * gn[q - 2];
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (krow) {
        const int f = j - 1, fc = max(f, 0), ip = min(i + 1, g.NX - 1);
        const T g0 = F(FNU, 2, i, fc, k), g1 = F(FNU, 2, ip, fc, k), g2 = F(FNT, 2, i, fc, k);
        if (f >= jl - 1 && f <= jh - 1) {  // V(i, j) advects u at (i, f) and (i+1, f) and theta at (i, f)
            if (i >= il && i < ih) gv += T(0.5) * g0;
            if (i + 1 >= il && i + 1 < ih) gv += T(0.5) * g1;
            if (icol) gv += g2;
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    // top faces kf = k - 2 .. k + 1 (q = kf - k + 2), unrolled like the loops above; kf = 1 and NZ - 2
    // are second-order faces
    if (icol && jrow) {
        T aw[4][3], sg[4][3];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const long long o = IX(i, j, min(max(k - 2 + q, 1), NZ - 2));
            aw[q][0] = ((const T *)fa.f[2 * FZU])[o];
            aw[q][1] = ((const T *)fa.f[2 * FZV])[o];
            aw[q][2] = ((const T *)fa.f[2 * FZT])[o];
            sg[q][0] = T(fa.sgn[FZU][o]);
            sg[q][1] = T(fa.sgn[FZV][o]);
            sg[q][2] = T(fa.sgn[FZT][o]);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int kf = k - 2 + q;
            if (kf < 1 || kf > NZ - 2) continue;
            if (kf == 1 || kf == NZ - 2) {
                if (k == kf || k == kf + 1) {
                    gu += T(0.5) * aw[q][0];
                    gv += T(0.5) * aw[q][1];
                    gth += T(0.5) * aw[q][2];
                }
            } else {
                const int d = 3 - q;  // d = k - kf + 1
                gu += aw[q][0] * (c4[d] + sg[q][0] * u3[d]);
                gv += aw[q][1] * (c4[d] + sg[q][1] * u3[d]);
                gth += aw[q][2] * (c4[d] + sg[q][2] * u3[d]);
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const int kf = k - 1;  // W(k) advects at the top face of level k-1
        const int kfc = min(max(kf, 0), NZ - 1), ip = min(i + 1, g.NX - 1), jp = min(j + 1, g.NY - 1);
        const int kc0 = min(max(k, 1), NZ - 1), kp = min(k + 1, NZ - 1);
        const T g0 = F(FZU, 2, i, j, kfc), g1 = F(FZU, 2, ip, j, kfc), g2 = F(FZV, 2, i, j, kfc), g3 = F(FZV, 2, i, jp, kfc);
        const T g4 = F(FZT, 2, i, j, kfc);
        const T ath0 = m.ATHc(i, j, k - 1), ath1 = m.ATHc(i, j, k), aw0 = m.AWc(i, j, k), aw1 = m.AWc(i, j, kp);
        const double rwk = c.rhow[kc0], tbk = c.tb[kc0], tbm = c.tb[kc0 - 1], odm = c.one_d_rhou[kc0 - 1], odk = c.one_d_rhou[kc0];
        const double tbvk = c.tbv[min(max(k, 0), NZ - 1)];
        if (kf >= 1 && kf <= NZ - 2) {
            if (jrow) {
                if (i >= il && i < ih) gw += T(0.5) * g0;
                if (i + 1 >= il && i + 1 < ih) gw += T(0.5) * g1;
            }
            if (icol) {
                if (j >= jl && j < jh) gw += T(0.5) * g2;
                if (j + 1 >= jl && j + 1 < jh) gw += T(0.5) * g3;
            }
            if (icol && jrow) gw += g4;
        }
        if (icol && jrow && k >= 1) {
            // Z_ADVECT_BASE (advection.cpp:16-18): W(k) enters the base-state term of levels k-1 and k
            const T coef = T(0.5) * T(rwk) * (T(tbk) - T(tbm));
            gw -= coef * (m.DTZ(k - 1) * T(odm) * ath0 + m.DTZ(k) * T(odk) * ath1);
        }
        // buoyancy (advection.cpp:329-337): WP(k') = WM + step dt g (th(k')/tbv(k') + th(k'-1)/tbv(k'-1)) / 2
        if (icol && jrow) gth += dt * T(kGrav) * T(0.5) * (aw0 + (k + 1 < NZ ? aw1 : T(0))) / T(tbvk);
    }
    const long long p = IX(i, j, k);
    a.CU[p] += gu;
    a.CV[p] += gv;
    a.CW[p] += gw;
    a.CTH[p] += gth;
}

// Pass 1 with every load unconditional: indices are pulled into the array (IXC) and only the stores are
// guarded, so a wavefront issues its loads together instead of waiting at each guarded block.  Capped at
// 96 VGPRs (five waves per SIMD; the uncapped form holds 120 and four), it measured 3 % faster than the
// guarded form on the 192x192x40 adjoint.  The same treatment of pass 2 as a whole needs 229 VGPRs (two
// waves per SIMD) and was slower; pass 2 batches its loads section by section instead, with a scheduling
// barrier between sections (83 VGPRs in FP64; 154 -> about 120 us at 192x192x40).
#define IXC(i, j, k) IX(min(max((i), 0), g.NX - 1), min(max((j), 0), g.NY - 1), min(max((k), 0), g.NZ - 1))

template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_adj_faces(Geo g, Cols c, AdvAdjArgs<T> a, AdjFaces fa, T step) {
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    const int NZ = g.NZ, il = a.il, ih = a.ih, jl = a.jl, jh = a.jh;
    auto L = [&](const T *F, int ii, int jj, int kk) -> T { return F[IXC(ii, jj, kk)]; };
    auto inb = [&](int ii, int jj, int kk) { return ii >= il && ii < ih && jj >= jl && jj < jh && kk >= 1 && kk <= NZ - 2; };
    auto A = [&](const T *P, int ii, int jj, int kk) -> T {
        const T v = step * L(P, ii, jj, kk);
        return inb(ii, jj, kk) ? v : T(0);
    };
    auto I5x = [&](const T *F, int e, int jj, int kk, T s) -> T {
        return i5(L(F, e - 2, jj, kk), L(F, e - 1, jj, kk), L(F, e, jj, kk), L(F, e + 1, jj, kk), L(F, e + 2, jj, kk), L(F, e + 3, jj, kk), s);
    };
    auto I5y = [&](const T *F, int ii, int f, int kk, T s) -> T {
        return i5(L(F, ii, f - 2, kk), L(F, ii, f - 1, kk), L(F, ii, f, kk), L(F, ii, f + 1, kk), L(F, ii, f + 2, kk), L(F, ii, f + 3, kk), s);
    };
    auto Iz = [&](const T *F, int ii, int jj, int kf, T s) -> T {
        const T lo = i2(L(F, ii, jj, kf), L(F, ii, jj, kf + 1));
        const T mid = i3(L(F, ii, jj, kf - 1), L(F, ii, jj, kf), L(F, ii, jj, kf + 1), L(F, ii, jj, kf + 2), s);
        return (kf == 1 || kf == NZ - 2) ? lo : mid;
    };
    const long long p = IX(i, j, k);
    auto put = [&](int kind, T fh, T wt, T gg) {
        ((T *)fa.f[2 * kind])[p] = fh * wt;
        ((T *)fa.f[2 * kind + 1])[p] = fh * gg;
        fa.sgn[kind][p] = wt > T(0) ? 1 : -1;
    };
    const T dtx = T(g.dtx), dty = T(g.dty);
    const bool krow = k >= 1 && k <= NZ - 2, jrow = j >= jl && j < jh, icol = i >= il && i < ih;
    const int kc = min(max(k, 0), NZ - 1);
    {  // east faces e = i
        const int e = i;
        const T ua = T(0.5) * (L(a.U, e + 1, j, k) + L(a.U, e, j, k)), uba = T(0.5) * (L(a.UB, e + 1, j, k) + L(a.UB, e, j, k));
        const T fu = dtx * (A(a.PU, e + 1, j, k) - A(a.PU, e, j, k));
        const T gu = I5x(a.UB, e, j, k, sgn(ua)) + I5x(a.U, e, j, k, sgn(uba + ua));
        const T va = T(0.5) * (L(a.U, e + 1, j, k) + L(a.U, e + 1, j - 1, k));
        const T vba = T(0.5) * (L(a.UB, e + 1, j, k) + L(a.UB, e + 1, j - 1, k));
        const T fv = dtx * (A(a.PV, e + 1, j, k) - A(a.PV, e, j, k));
        const T gv = I5x(a.VB, e, j, k, sgn(va)) + I5x(a.V, e, j, k, sgn(vba + va));
        const T U1 = L(a.U, e + 1, j, k), UB1 = L(a.UB, e + 1, j, k);
        const T ft = dtx * (A(a.PTH, e + 1, j, k) - A(a.PTH, e, j, k));
        const T gt = I5x(a.THB, e, j, k, sgn(U1)) + I5x(a.TH, e, j, k, sgn(UB1 + U1));
        if (i >= il - 1 && i <= ih - 1 && jrow && krow) {
            put(FEU, fu, uba + ua, gu);
            put(FEV, fv, vba + va, gv);
            put(FET, ft, UB1 + U1, gt);
        }
    }
    {  // north faces f = j
        const int f = j;
        const T va = T(0.5) * (L(a.V, i, f + 1, k) + L(a.V, i - 1, f + 1, k));
        const T vba = T(0.5) * (L(a.VB, i, f + 1, k) + L(a.VB, i - 1, f + 1, k));
        const T fu = dty * (A(a.PU, i, f + 1, k) - A(a.PU, i, f, k));
        const T gu = I5y(a.UB, i, f, k, sgn(va)) + I5y(a.U, i, f, k, sgn(vba + va));
        const T wa = T(0.5) * (L(a.V, i, f + 1, k) + L(a.V, i, f, k)), wba = T(0.5) * (L(a.VB, i, f + 1, k) + L(a.VB, i, f, k));
        const T fv = dty * (A(a.PV, i, f + 1, k) - A(a.PV, i, f, k));
        const T gv = I5y(a.VB, i, f, k, sgn(wa)) + I5y(a.V, i, f, k, sgn(wba + wa));
        const T V1 = L(a.V, i, f + 1, k), VB1 = L(a.VB, i, f + 1, k);
        const T ft = dty * (A(a.PTH, i, f + 1, k) - A(a.PTH, i, f, k));
        const T gt = I5y(a.THB, i, f, k, sgn(V1)) + I5y(a.TH, i, f, k, sgn(VB1 + V1));
        if (j >= jl - 1 && j <= jh - 1 && icol && krow) {
            put(FNU, fu, vba + va, gu);
            put(FNV, fv, wba + wa, gv);
            put(FNT, ft, VB1 + V1, gt);
        }
    }
    {  // top faces kf = k
        const int kf = k;
        const T zlo = -T(g.dtz) * T(c.mu[kc]) * T(c.rhow[min(kc + 1, NZ - 1)]) * T(c.one_d_rhou[kc]);
        const T zhi = T(g.dtz) * T(c.mu[min(kc + 1, NZ - 1)]) * T(c.rhow[min(kc + 1, NZ - 1)]) * T(c.one_d_rhou[min(kc + 1, NZ - 1)]);
        auto zf = [&](T A0, T A1) { return zlo * A0 + zhi * A1; };
        const T wa = T(0.5) * (L(a.W, i, j, kf + 1) + L(a.W, i - 1, j, kf + 1));
        const T wba = T(0.5) * (L(a.WB, i, j, kf + 1) + L(a.WB, i - 1, j, kf + 1));
        const T fu = zf(A(a.PU, i, j, kf), A(a.PU, i, j, kf + 1));
        const T gu = Iz(a.UB, i, j, kf, sgn(wa)) + Iz(a.U, i, j, kf, sgn(wba + wa));
        const T xa = T(0.5) * (L(a.W, i, j, kf + 1) + L(a.W, i, j - 1, kf + 1));
        const T xba = T(0.5) * (L(a.WB, i, j, kf + 1) + L(a.WB, i, j - 1, kf + 1));
        const T fv = zf(A(a.PV, i, j, kf), A(a.PV, i, j, kf + 1));
        const T gv = Iz(a.VB, i, j, kf, sgn(xa)) + Iz(a.V, i, j, kf, sgn(xba + xa));
        const T W1 = L(a.W, i, j, kf + 1), WB1 = L(a.WB, i, j, kf + 1);
        const T ft = zf(A(a.PTH, i, j, kf), A(a.PTH, i, j, kf + 1));
        const T gt = Iz(a.THB, i, j, kf, sgn(W1)) + Iz(a.TH, i, j, kf, sgn(WB1 + W1));
        if (krow && icol && jrow) {
            put(FZU, fu, wba + wa, gu);
            put(FZV, fv, xba + xa, gv);
            put(FZT, ft, WB1 + W1, gt);
        }
    }
}

#undef IXC

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
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
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

// up to 12 equal-sized device copies in one launch (grid.y = copy index), 16 bytes per lane and
// 4-byte words for the tail; `words` = bytes / 4
struct CopyList {
    const void *src[12];
    void *dst[12];
};
__global__ void __launch_bounds__(256) k_copy_list(CopyList cl, long long words) {
    const uint4 *s = (const uint4 *)cl.src[blockIdx.y];
    uint4 *d = (uint4 *)cl.dst[blockIdx.y];
    const long long n16 = words / 4;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n16; q += (long long)gridDim.x * blockDim.x) d[q] = s[q];
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)(words - n16 * 4))
        ((unsigned *)d)[n16 * 4 + threadIdx.x] = ((const unsigned *)s)[n16 * 4 + threadIdx.x];
}
int launch_copy_list(Ctx *c, const CopyList &cl, int n, size_t bytes) {
    for (int q = 0; q < n; q++)
        if (((uintptr_t)cl.src[q] | (uintptr_t)cl.dst[q]) % 16 != 0 || bytes % 4 != 0) {
            set_error("adjoint copy: buffers must be 16-byte aligned and sizes whole words");
            return DYN_E_ARG;
        }
    const long long words = (long long)(bytes / 4);
    const long long blocks = std::min<long long>((words / 4 + 255) / 256, 2048);
    hipLaunchKernelGGL(k_copy_list, dim3((unsigned)std::max<long long>(blocks, 1), (unsigned)n), dim3(256), 0, c->stream, cl, words);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

}  // namespace

// Trajectory storage and adjoint buffers of one context (allocated by dyn_adjoint_record).
struct AdjState {
    int nsteps = 0, every = 0;
    size_t field_bytes = 0;
    std::vector<void *> snap;  // snapshot s, field f, role r at snap[(s * 4 + f) * 3 + r]; last = the final state
    // two records of a segment's stage states: step n, stage st, field f at rec[b][(n * 3 + st) * 4 + f];
    // the second is allocated by the reverse sweep when there is more than one segment, so that the
    // next segment down is recomputed into one while the adjoint steps read the other
    std::vector<void *> rec[2];
    std::vector<void *> slabs;   // allocations snap and rec are carved from
    int rec_seg[2] = {-1, -1};   // segment whose stage states rec[b] holds (-1: none)
    int rec_w = 0, rec_r = 0;    // record written by the stage hook / read by the adjoint steps
    // the reverse sweep's adjoint steps run on their own stream with their own spectrum scratch
    hipStream_t astream = nullptr;
    hipEvent_t ev_rec[2] = {nullptr, nullptr}, ev_adj[2] = {nullptr, nullptr};
    bool adj_pending[2] = {false, false};  // ev_adj[b]: adjoint steps that read rec[b] were queued
    void *spec_a = nullptr, *spec2_a = nullptr;
    void *adj[kAdjFields][3] = {{nullptr}};
    void *pihat = nullptr, *r1 = nullptr, *r2 = nullptr;
    void *faces[18] = {nullptr};  // per-face adjoint weights of the advection transpose (Aw, G per kind)
    void *face_sgn[9] = {nullptr};  // per-face upwind signs (one byte each)
    double *stage_dbl = nullptr;  // host-transfer staging (double)
    int rec_base = 0;             // first recorded step of the segment in the record
};

namespace {

void adj_free(Ctx *c) {
    AdjState *s = c->adj;
    if (!s) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : s->slabs) (void)hipFree(p);
    for (auto &f : s->adj)
        for (void *p : f)
            if (p) (void)hipFree(p);
    if (s->astream) (void)hipStreamSynchronize(s->astream);
    for (hipEvent_t e : {s->ev_rec[0], s->ev_rec[1], s->ev_adj[0], s->ev_adj[1]})
        if (e) (void)hipEventDestroy(e);
    if (s->astream) (void)hipStreamDestroy(s->astream);
    void *ptrs[] = {s->pihat, s->r1, s->r2, s->stage_dbl, s->spec_a, s->spec2_a};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (void *p : s->faces)
        if (p) (void)hipFree(p);
    for (void *p : s->face_sgn)
        if (p) (void)hipFree(p);
    delete s;
    c->adj = nullptr;
}

// `count` fields of `bytes` each carved from one allocation, 256-byte aligned
int alloc_fields(AdjState *s, std::vector<void *> &v, size_t count, size_t bytes) {
    const size_t stride = (bytes + 255) / 256 * 256;
    void *slab = nullptr;
    if (hipMalloc(&slab, stride * count) != hipSuccess) {
        (void)hipGetLastError();
        set_error("dyn_adjoint_record: not enough device memory for the trajectory (" + std::to_string(stride * count >> 20) +
                  " MiB); use a shorter checkpoint interval");
        return DYN_E_HIP;
    }
    s->slabs.push_back(slab);
    v.resize(count);
    for (size_t q = 0; q < count; q++) v[q] = (char *)slab + q * stride;
    return DYN_OK;
}

int copy_state(Ctx *c, void *const *dst_or_null, void *const *src_or_null, bool to_snapshot) {
    AdjState *s = c->adj;
    CopyList cl{};
    for (int f = 0; f < kAdjFields; f++)
        for (int r = 0; r < 3; r++) {
            void *prog = c->prog[kAdjIds[f]][r];
            void *snap = to_snapshot ? dst_or_null[f * 3 + r] : src_or_null[f * 3 + r];
            cl.src[f * 3 + r] = to_snapshot ? prog : snap;
            cl.dst[f * 3 + r] = to_snapshot ? snap : prog;
        }
    return launch_copy_list(c, cl, kAdjFields * 3, s->field_bytes);
}

// stage hook of the recomputation: the current buffers at the start of every stage
int record_hook(Ctx *c, int stage, void *arg) {
    const int n = *(const int *)arg;
    AdjState *s = c->adj;
    CopyList cl{};
    for (int f = 0; f < kAdjFields; f++) {
        cl.src[f] = c->prog[kAdjIds[f]][DYN_ROLE_CUR];
        cl.dst[f] = s->rec[s->rec_w][((size_t)n * 3 + stage) * 4 + f];
    }
    return launch_copy_list(c, cl, kAdjFields, s->field_bytes);
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
        hipLaunchKernelGGL(k_adj_pgf<T>, point_grid(g), dim3(256), 0, c->stream, g, c->cols,
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
    const std::vector<void *> &rec = s->rec[s->rec_r];
    a.U = (const T *)rec[rb + 0];
    a.V = (const T *)rec[rb + 1];
    a.W = (const T *)rec[rb + 2];
    a.TH = (const T *)rec[rb + 3];
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
        AdjFaces fa;
        static_assert(sizeof(fa.f) == sizeof(s->faces) && sizeof(fa.sgn) == sizeof(s->face_sgn), "face array count");
        memcpy(fa.f, s->faces, sizeof(fa.f));
        for (int k = 0; k < NFACE; k++) fa.sgn[k] = (signed char *)s->face_sgn[k];
        Timed _t(c, "adj_advect", (double)c->npts);
        hipLaunchKernelGGL(k_adj_faces<T>, point_grid(g), dim3(256), 0, c->stream, g, c->cols, a, fa, T(step));
        DYN_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_adj_gather<T>, point_grid(g), dim3(256), 0, c->stream, g, c->cols, a, fa, T(step));
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

// the second record, the adjoint stream, its events and its spectrum scratch (once per record pass); on
// failure the sweep runs serially
int ensure_pipeline(Ctx *c) {
    AdjState *s = c->adj;
    if (s->astream) return DYN_OK;
    const size_t spec_bytes = sizeof(double) * 2 * (size_t)c->pp.ni * c->pp.nj * c->pp.nk2;
    if (alloc_fields(s, s->rec[1], (size_t)s->every * 12, s->field_bytes) != DYN_OK) {
        set_error("");
        return DYN_E_HIP;
    }
    if (hipMalloc(&s->spec_a, spec_bytes) != hipSuccess || (c->pp.spec2 && hipMalloc(&s->spec2_a, spec_bytes) != hipSuccess)) {
        (void)hipGetLastError();
        return DYN_E_HIP;
    }
    for (hipEvent_t *e : {&s->ev_rec[0], &s->ev_rec[1], &s->ev_adj[0], &s->ev_adj[1]})
        DYN_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    DYN_HIP(hipStreamCreate(&s->astream));
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
    // recompute segment seg from its snapshot into record b (on the context's streams)
    auto recompute = [&](int seg, int b) -> int {
        const int n0 = seg * s->every, n1 = std::min(s->nsteps, n0 + s->every);
        int rc = copy_state(c, nullptr, &s->snap[(size_t)seg * 12], false);
        if (rc) return rc;
        s->rec_w = b;
        for (int n = n0; n < n1; n++) {
            int local = n - n0;
            rc = step_with_hook(c, record_hook, &local, n < n1 - 1);
            if (rc) {
                unshare_roles(c);
                return rc;
            }
        }
        s->rec_seg[b] = seg;
        return DYN_OK;
    };
    // the adjoint steps of segment seg from record b, on `stream` with spectrum scratch spec/spec2
    auto adjoint_steps = [&](int seg, int b, hipStream_t stream, void *spec, void *spec2) -> int {
        const int n0 = seg * s->every, n1 = std::min(s->nsteps, n0 + s->every);
        hipStream_t own = c->stream;
        void *sp = c->pp.spec, *sp2 = c->pp.spec2;
        c->stream = stream;
        c->pp.spec = spec;
        c->pp.spec2 = spec2;
        s->rec_r = b;
        s->rec_base = n0;
        int rc = DYN_OK;
        for (int n = n1 - 1; n >= n0 && !rc; n--) rc = step_adjoint<T>(c, n);
        c->stream = own;
        c->pp.spec = sp;
        c->pp.spec2 = sp2;
        return rc;
    };
    const bool pipe = nseg > 1 && ensure_pipeline(c) == DYN_OK;
    if (!pipe) {
        for (int seg = nseg - 1; seg >= 0; seg--) {
            int rc = DYN_OK;
            if (s->rec_seg[0] != seg) rc = recompute(seg, 0);  // the record pass keeps the last segment
            if (!rc) rc = adjoint_steps(seg, 0, c->stream, c->pp.spec, c->pp.spec2);
            if (rc) return rc;
        }
    } else {
        // Segment seg's adjoint steps run on the adjoint stream while segment seg-1 is recomputed on the
        // context's stream into the other record; events order each record's writes and reads.  Beside
        // the adjoint steps the recompute runs each stage on the context's stream alone: a second stream
        // of its own only adds to what competes with the adjoint stream (the sweep measured 1 % faster
        // that way).  The stream knob is put back on every return.
        struct OneStream {
            Ctx *c;
            int saved;
            explicit OneStream(Ctx *cc) : c(cc), saved(cc->tune[DYN_TUNE_STAGE_STREAMS]) { c->tune[DYN_TUNE_STAGE_STREAMS] = 1; }
            ~OneStream() { c->tune[DYN_TUNE_STAGE_STREAMS] = saved; }
        } one_stream(c);
        DYN_HIP(hipEventRecord(s->ev_rec[0], c->stream));  // the seed and the record pass come first
        DYN_HIP(hipStreamWaitEvent(s->astream, s->ev_rec[0], 0));
        s->adj_pending[0] = s->adj_pending[1] = false;
        int b = -1;
        for (int seg = nseg - 1; seg >= 0; seg--) {
            b = s->rec_seg[0] == seg ? 0 : (s->rec_seg[1] == seg ? 1 : -1);
            if (b < 0) {  // not recorded (a second sweep over one record pass)
                b = 0;
                if (s->adj_pending[b]) DYN_HIP(hipStreamWaitEvent(c->stream, s->ev_adj[b], 0));
                int rc = recompute(seg, b);
                if (rc) return rc;
                DYN_HIP(hipEventRecord(s->ev_rec[b], c->stream));
            }
            if (seg > 0) {
                const int ob = 1 - b;
                if (s->adj_pending[ob]) DYN_HIP(hipStreamWaitEvent(c->stream, s->ev_adj[ob], 0));
                int rc = recompute(seg - 1, ob);
                if (rc) return rc;
                DYN_HIP(hipEventRecord(s->ev_rec[ob], c->stream));
            }
            DYN_HIP(hipStreamWaitEvent(s->astream, s->ev_rec[b], 0));
            int rc = adjoint_steps(seg, b, s->astream, s->spec_a, s->spec2_a);
            if (rc) return rc;
            DYN_HIP(hipEventRecord(s->ev_adj[b], s->astream));
            s->adj_pending[b] = true;
        }
        DYN_HIP(hipStreamWaitEvent(c->stream, s->ev_adj[b], 0));  // the gradients follow the last adjoint step
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
    if (c->turb.on) {
        set_error("dyn_adjoint_record: the adjoint covers the step without turbulence");
        return DYN_E_UNSUPPORTED;
    }
    if (nsteps <= 0 || every < 0) {
        set_error("dyn_adjoint_record: nsteps must be positive and every non-negative");
        return DYN_E_ARG;
    }
    if (c->npts >= (size_t(1) << 31)) {
        set_error("dyn_adjoint_record: the adjoint kernels take fields of fewer than 2^31 points");
        return DYN_E_UNSUPPORTED;
    }
    adj_free(c);
    AdjState *s = new AdjState();
    c->adj = s;
    s->nsteps = nsteps;
    s->field_bytes = c->elt * c->npts;
    const size_t fstride = (s->field_bytes + 255) / 256 * 256;
    if (every == 0) {
        // longest segment whose records (12 fields a step) and snapshots (12 fields each) fit in half of
        // the free device memory, less the ~40 fields of adjoint and face buffers allocated below
        size_t freeb = 0, totalb = 0;
        DYN_HIP(hipMemGetInfo(&freeb, &totalb));
        const double budget = 0.5 * (double)freeb - 40.0 * (double)fstride;
        every = nsteps;
        while (every > 1) {
            const int nseg = (nsteps + every - 1) / every;
            if ((double)(every + nseg + 1) * 12.0 * (double)fstride <= budget) break;
            every = (every + 1) / 2;
        }
    }
    s->every = std::min(every, nsteps);
    const int nseg = (nsteps + s->every - 1) / s->every;
    int rc = alloc_fields(s, s->snap, (size_t)(nseg + 1) * 12, s->field_bytes);
    if (!rc) rc = alloc_fields(s, s->rec[0], (size_t)s->every * 12, s->field_bytes);
    if (rc) {
        adj_free(c);
        return rc;
    }
    for (auto &f : s->adj)
        for (auto &p : f) DYN_HIP(hipMalloc(&p, s->field_bytes));
    DYN_HIP(hipMalloc(&s->pihat, s->field_bytes));
    DYN_HIP(hipMalloc(&s->r1, s->field_bytes));
    DYN_HIP(hipMalloc(&s->r2, s->field_bytes));
    for (auto &p : s->faces) DYN_HIP(hipMalloc(&p, s->field_bytes));
    for (auto &p : s->face_sgn) DYN_HIP(hipMalloc(&p, c->npts));
    DYN_HIP(hipMalloc(&s->stage_dbl, sizeof(double) * c->npts));
    // forward pass with a snapshot every `every` steps and one of the final state; the last segment's
    // stage states are recorded on the way, so the reverse sweep starts without a recomputation
    const int last0 = (nseg - 1) * s->every;
    for (int n = 0; n < nsteps;) {
        if (n % s->every == 0) {
            rc = copy_state(c, &s->snap[(size_t)(n / s->every) * 12], nullptr, true);
            if (rc) return rc;
        }
        if (n < last0) {
            // a segment before the last records nothing: its steps replay as dyn_step's graphs (the
            // same steps bit for bit), which leave the roles apart for the next snapshot
            rc = dyn_step(reinterpret_cast<dyn_ctx *>(c), s->every);
            if (rc) return rc;
            n += s->every;
            continue;
        }
        int local = n - last0;
        // M and current share a buffer between the steps (the snapshots copy each role's values, the
        // same whether or not they share); the last step separates them
        const bool share = n < nsteps - 1;
        rc = step_with_hook(c, record_hook, &local, share);
        if (rc) {
            unshare_roles(c);
            return rc;
        }
        n++;
    }
    s->rec_seg[0] = nseg - 1;
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

int dyn_adjoint_interval(dyn_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) {
        set_error("null context");
        return DYN_E_ARG;
    }
    if (!c->adj) {
        set_error("dyn_adjoint_interval: no recorded trajectory");
        return DYN_E_STATE;
    }
    return c->adj->every;
}

int dyn_adjoint_release(dyn_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DYN_E_ARG;
    adj_free(c);
    return DYN_OK;
}

}  // extern "C"
