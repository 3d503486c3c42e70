// Fused stage advection for gfx950: advect_uvw_velocity + advect_theta + advect_microphysics_cell
// (+ the fall speed and the interior part of zero_moisture) in one pass over HBM.
//
// Mapping: one 64-lane wavefront per j-row, lane = level k (NZ <= 64); a block of TJ rows marches
// along i over a chunk of TI columns, as the reference's slab loop does (advection.cpp:313-325):
//   * i-stencils come from per-thread register queues (own j,k; columns i-3..i+3) and the east-face
//     flux of column i is carried as the west face of column i+1 (fluxes.cpp:891,1322,1504);
//   * j- and k-stencils read LDS planes of TJ+6 rows; U/UB are kept for planes i and i+1, V/VB/W/WB for
//     planes i-1 and i (staggered advecting velocities), the scalars for plane i;
//   * north and top faces are computed once per point and exchanged through LDS as the south and
//     bottom faces of the neighbours.
// Every face and tendency uses the same expression tree as the per-operator kernels, so the fused and
// unfused paths agree to the last bit.
#include "dyn_internal.hpp"

namespace dyn {
namespace {

constexpr int KP = 64;  // lanes per row (levels)

template <typename T>
__device__ __forceinline__ T sgn(T x) { return x > T(0) ? T(1) : T(-1); }

// interpolants of fluxes.cpp:110-206 on explicit operands (p0 = value on the low side of the face)
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

// queue fields (register, own j,k, columns i-3..i+3)
enum { QU = 0, QUB, QV, QVB, QTH, QTHB, QQV, QQB, QQC, QQR, NQ };
// LDS plane fields
enum { LU = 0, LUB, LV, LVB, LW, LWB, LTH, LTHB, LQV, LQB, LQC, LQR, NL };
// face quantities
enum { FU = 0, FV, FTH, FQV, FQC, FQR, NF };

struct FusedArgs {
    const void *cur[7];   // U V W TH QV QC QR (current)
    const void *mprev[7]; // M buffers
    void *out[7];         // P buffers
    const void *bar[5];   // UB VB WB THB QB
    void *vt;
    int il, ih, jl, jh;   // advected box
    int ti;               // columns per block
};

template <typename T, int TJ, bool MOIST>
struct Smem {
    static constexpr int R = TJ + 6;
    // ring of 3 planes for U,UB,V,VB,W,WB and one plane for the scalars
    static constexpr int NWIN = 6;
    static constexpr int NSCAL = MOIST ? 6 : 2;
    static constexpr size_t plane = (size_t)R * KP;
    static constexpr size_t win_elems = 3 * NWIN * plane;
    static constexpr size_t scal_elems = NSCAL * plane;
    static constexpr size_t north_elems = NF * (TJ + 1) * KP;
    static constexpr size_t top_elems = NF * TJ * KP;
    static constexpr size_t vt_elems = TJ * KP;
    static constexpr size_t total = win_elems + scal_elems + north_elems + top_elems + vt_elems;
    static constexpr size_t bytes = total * sizeof(T);
};

template <typename T, int TJ, bool MOIST>
__global__ void __launch_bounds__(KP *TJ) k_stage_advect(Geo g, Cols c, FusedArgs a, T step) {
    using S = Smem<T, TJ, MOIST>;
    extern __shared__ unsigned char smem_raw[];
    T *sm = (T *)smem_raw;
    T *win = sm;                       // [slot][field][row][k]
    T *scal = win + S::win_elems;      // [field][row][k]
    T *nface = scal + S::scal_elems;   // [face][row+1][k]
    T *tface = nface + S::north_elems; // [face][row][k]
    T *vtp = tface + S::top_elems;     // [row][k]

    const int k = threadIdx.x;
    const int r = threadIdx.y;
    const int j0 = a.jl + blockIdx.x * TJ;
    const int j = j0 + r;
    const int i0 = a.il + blockIdx.y * a.ti;
    const int i1 = min(i0 + a.ti, a.ih);
    if (i0 >= i1) return;
    const int NZ = g.NZ;
    const bool kin = k < NZ;
    const bool row_ok = j < a.jh;
    const bool active = row_ok && k >= 1 && k <= NZ - 2;
    const int kc = kin ? k : NZ - 1;  // clamped level for loads
    const int jc = min(j, g.NY - 1);

    const T *U = (const T *)a.cur[0], *V = (const T *)a.cur[1], *W = (const T *)a.cur[2], *TH = (const T *)a.cur[3];
    const T *QVf = (const T *)a.cur[4], *QCf = (const T *)a.cur[5], *QRf = (const T *)a.cur[6];
    const T *UB = (const T *)a.bar[0], *VB = (const T *)a.bar[1], *WB = (const T *)a.bar[2], *THB = (const T *)a.bar[3];
    const T *QB = (const T *)a.bar[4];
    const T *winsrc[6] = {U, UB, V, VB, W, WB};
    const T *scalsrc[6] = {TH, THB, QVf, QB, QCf, QRf};
    const T *qsrc[NQ] = {U, UB, V, VB, TH, THB, QVf, QB, QCf, QRf};
    constexpr int nq = MOIST ? NQ : 6;

    auto W_AT = [&](int slot, int f, int row, int kk) -> T & { return win[((slot * 6 + f) * S::R + row) * KP + kk]; };
    auto S_AT = [&](int f, int row, int kk) -> T & { return scal[(f * S::R + row) * KP + kk]; };
    auto slot_of = [](int i) { return ((i % 3) + 3) % 3; };
    auto gidx = [&](int i, int jj, int kk) -> long long { return (long long)i * g.NYNZ + (long long)jj * NZ + kk; };

    // cooperative load of one plane of a set of fields (rows j0-3 .. j0+TJ+2)
    auto load_win_plane = [&](int i) {
        int slot = slot_of(i);
        for (int rr = r; rr < S::R; rr += TJ) {
            int jj = min(max(j0 - 3 + rr, 0), g.NY - 1);
            long long p = gidx(i, jj, kc);
#pragma unroll
            for (int f = 0; f < 6; f++) W_AT(slot, f, rr, k) = winsrc[f][p];
        }
    };
    auto load_scal_plane = [&](int i) {
        for (int rr = r; rr < S::R; rr += TJ) {
            int jj = min(max(j0 - 3 + rr, 0), g.NY - 1);
            long long p = gidx(i, jj, kc);
#pragma unroll
            for (int f = 0; f < S::NSCAL; f++) S_AT(f, rr, k) = scalsrc[f][p];
        }
    };

    // register queues: q[f][d] holds column i-3+d (d = 0..6) at iteration i
    T q[NQ][7];
#pragma unroll
    for (int f = 0; f < nq; f++)
#pragma unroll
        for (int d = 0; d < 7; d++) q[f][d] = qsrc[f][gidx(i0 - 3 + d, jc, kc)];

    load_win_plane(i0 - 1);
    load_win_plane(i0);
    load_win_plane(i0 + 1);

    const T dtx = T(g.dtx), dty = T(g.dty), dt = T(g.dt);
    const T DTZ = T(g.dtz) * T(c.mu[kc]);
    const T rw1 = T(c.rhow[min(kc + 1, NZ - 1)]), rw0 = T(c.rhow[kc]), odr = T(c.one_d_rhou[kc]);
    const T fj = T(c.f[jc]);
    const int R = r + 3;

    // east-face fluxes of column i (queue positions: column i at d=3 after the shift below)
    auto east_faces = [&](int base, T *e) {
        // base: queue index of column "i" (the low side of the face i+1/2)
#define QQ(f, o) q[f][base + (o)]
        {  // u, u_east
            T ua = T(0.5) * (QQ(QU, 1) + QQ(QU, 0));
            T uba = T(0.5) * (QQ(QUB, 1) + QQ(QUB, 0));
            T ue = i5(QQ(QU, -2), QQ(QU, -1), QQ(QU, 0), QQ(QU, 1), QQ(QU, 2), QQ(QU, 3), sgn(uba + ua));
            T ube = i5(QQ(QUB, -2), QQ(QUB, -1), QQ(QUB, 0), QQ(QUB, 1), QQ(QUB, 2), QQ(QUB, 3), sgn(ua));
            e[FU] = ua * ube + (uba + ua) * ue;
        }
        // th (s_east)
        {
            T U1 = QQ(QU, 1), UB1 = QQ(QUB, 1);
            T se = i5(QQ(QTH, -2), QQ(QTH, -1), QQ(QTH, 0), QQ(QTH, 1), QQ(QTH, 2), QQ(QTH, 3), sgn(UB1 + U1));
            T be = i5(QQ(QTHB, -2), QQ(QTHB, -1), QQ(QTHB, 0), QQ(QTHB, 1), QQ(QTHB, 2), QQ(QTHB, 3), sgn(U1));
            e[FTH] = U1 * be + (UB1 + U1) * se;
            if (MOIST) {
                T sv = i5(QQ(QQV, -2), QQ(QQV, -1), QQ(QQV, 0), QQ(QQV, 1), QQ(QQV, 2), QQ(QQV, 3), sgn(UB1 + U1));
                T bv = i5(QQ(QQB, -2), QQ(QQB, -1), QQ(QQB, 0), QQ(QQB, 1), QQ(QQB, 2), QQ(QQB, 3), sgn(U1));
                e[FQV] = U1 * bv + (UB1 + U1) * sv;
                T ub = UB1 + U1;
                e[FQC] = ub * i5(QQ(QQC, -2), QQ(QQC, -1), QQ(QQC, 0), QQ(QQC, 1), QQ(QQC, 2), QQ(QQC, 3), sgn(ub));
                e[FQR] = ub * i5(QQ(QQR, -2), QQ(QQR, -1), QQ(QQR, 0), QQ(QQR, 1), QQ(QQR, 2), QQ(QQR, 3), sgn(ub));
            }
        }
#undef QQ
    };
    // v east face needs U at column i+1 and row j-1 (window plane i+1)
    auto v_east_face = [&](int base, int i) -> T {
        int sp = slot_of(i + 1);
        T ua = T(0.5) * (q[QU][base + 1] + W_AT(sp, 0, R - 1, k));
        T uba = T(0.5) * (q[QUB][base + 1] + W_AT(sp, 1, R - 1, k));
        T ve = i5(q[QV][base - 2], q[QV][base - 1], q[QV][base], q[QV][base + 1], q[QV][base + 2], q[QV][base + 3], sgn(uba + ua));
        T vbe = i5(q[QVB][base - 2], q[QVB][base - 1], q[QVB][base], q[QVB][base + 1], q[QVB][base + 2], q[QVB][base + 3], sgn(ua));
        return ua * vbe + (uba + ua) * ve;
    };

    T west[NF], east[NF];
    // west faces of the first column: the east faces of column i0-1 (compute_west*)
    __syncthreads();
    east_faces(2, west);
    west[FV] = v_east_face(2, i0 - 1);

    for (int i = i0; i < i1; i++) {
        const int s0 = slot_of(i), sm1 = slot_of(i - 1);
        load_scal_plane(i);
        __syncthreads();
        // ---- faces of column i -----------------------------------------------------------------
        east_faces(3, east);
        east[FV] = v_east_face(3, i);
        // north faces for rows r (and r-1 for the first row): rows of plane i
        auto north_faces = [&](int RR, T *nf) {
            {  // u_north
                T va = T(0.5) * (W_AT(s0, 2, RR + 1, k) + W_AT(sm1, 2, RR + 1, k));
                T vba = T(0.5) * (W_AT(s0, 3, RR + 1, k) + W_AT(sm1, 3, RR + 1, k));
                T un = i5(W_AT(s0, 0, RR - 2, k), W_AT(s0, 0, RR - 1, k), W_AT(s0, 0, RR, k), W_AT(s0, 0, RR + 1, k),
                          W_AT(s0, 0, RR + 2, k), W_AT(s0, 0, RR + 3, k), sgn(vba + va));
                T ubn = i5(W_AT(s0, 1, RR - 2, k), W_AT(s0, 1, RR - 1, k), W_AT(s0, 1, RR, k), W_AT(s0, 1, RR + 1, k),
                           W_AT(s0, 1, RR + 2, k), W_AT(s0, 1, RR + 3, k), sgn(va));
                nf[FU] = va * ubn + (vba + va) * un;
            }
            {  // v_north
                T va = T(0.5) * (W_AT(s0, 2, RR + 1, k) + W_AT(s0, 2, RR, k));
                T vba = T(0.5) * (W_AT(s0, 3, RR + 1, k) + W_AT(s0, 3, RR, k));
                T vn = i5(W_AT(s0, 2, RR - 2, k), W_AT(s0, 2, RR - 1, k), W_AT(s0, 2, RR, k), W_AT(s0, 2, RR + 1, k),
                          W_AT(s0, 2, RR + 2, k), W_AT(s0, 2, RR + 3, k), sgn(vba + va));
                T vbn = i5(W_AT(s0, 3, RR - 2, k), W_AT(s0, 3, RR - 1, k), W_AT(s0, 3, RR, k), W_AT(s0, 3, RR + 1, k),
                           W_AT(s0, 3, RR + 2, k), W_AT(s0, 3, RR + 3, k), sgn(va));
                nf[FV] = va * vbn + (vba + va) * vn;
            }
            T V1 = W_AT(s0, 2, RR + 1, k), VB1 = W_AT(s0, 3, RR + 1, k);
            {  // s_north (theta)
                T sn = i5(S_AT(0, RR - 2, k), S_AT(0, RR - 1, k), S_AT(0, RR, k), S_AT(0, RR + 1, k), S_AT(0, RR + 2, k),
                          S_AT(0, RR + 3, k), sgn(VB1 + V1));
                T bn = i5(S_AT(1, RR - 2, k), S_AT(1, RR - 1, k), S_AT(1, RR, k), S_AT(1, RR + 1, k), S_AT(1, RR + 2, k),
                          S_AT(1, RR + 3, k), sgn(V1));
                nf[FTH] = V1 * bn + (VB1 + V1) * sn;
            }
            if (MOIST) {
                T sn = i5(S_AT(2, RR - 2, k), S_AT(2, RR - 1, k), S_AT(2, RR, k), S_AT(2, RR + 1, k), S_AT(2, RR + 2, k),
                          S_AT(2, RR + 3, k), sgn(VB1 + V1));
                T bn = i5(S_AT(3, RR - 2, k), S_AT(3, RR - 1, k), S_AT(3, RR, k), S_AT(3, RR + 1, k), S_AT(3, RR + 2, k),
                          S_AT(3, RR + 3, k), sgn(V1));
                nf[FQV] = V1 * bn + (VB1 + V1) * sn;
                T vb = VB1 + V1;
                nf[FQC] = vb * i5(S_AT(4, RR - 2, k), S_AT(4, RR - 1, k), S_AT(4, RR, k), S_AT(4, RR + 1, k), S_AT(4, RR + 2, k),
                                  S_AT(4, RR + 3, k), sgn(vb));
                nf[FQR] = vb * i5(S_AT(5, RR - 2, k), S_AT(5, RR - 1, k), S_AT(5, RR, k), S_AT(5, RR + 1, k), S_AT(5, RR + 2, k),
                                  S_AT(5, RR + 3, k), sgn(vb));
            }
        };
        T nf[NF];
        north_faces(R, nf);
#pragma unroll
        for (int f = 0; f < NF; f++) nface[(f * (TJ + 1) + r + 1) * KP + k] = nf[f];
        if (r == 0) {
            T sf[NF];
            north_faces(R - 1, sf);
#pragma unroll
            for (int f = 0; f < NF; f++) nface[(f * (TJ + 1)) * KP + k] = sf[f];
        }
        // fall speed of this level (kessler.cpp:222-252); level 0 copies level 1, NZ-1 comes from the array
        T vtk = T(0);
        if (MOIST) {
            if (k >= 1 && k <= NZ - 2) {
                T qr = S_AT(5, R, k);
                if (qr > T(1.0e-12)) {
                    T rk = T(c.rhou[k]);
                    T qrr = fmax(qr * T(0.001) * rk, T(0));
                    T vtden = sqrt(T(c.rhou[1]) / rk);
                    vtk = T(36.34) * (pow(qrr, T(0.1364))) * vtden;
                }
            } else if (k == NZ - 1) {
                vtk = ((const T *)a.vt)[gidx(i, jc, k)];
            }
            vtp[r * KP + k] = vtk;
        }
        // top faces of this level (k+1/2); level 0 has none
        T tf[NF];
        {
            const int kp1 = min(kc + 1, NZ - 1), km1 = max(kc - 1, 0), kp2 = min(kc + 2, NZ - 1);
            const bool low = (k == 1 || k == NZ - 2);
            {  // u_top
                T wa = T(0.5) * (W_AT(s0, 4, R, kp1) + W_AT(sm1, 4, R, kp1));
                T wba = T(0.5) * (W_AT(s0, 5, R, kp1) + W_AT(sm1, 5, R, kp1));
                T ut, ubt;
                if (low) {
                    ut = i2(W_AT(s0, 0, R, kc), W_AT(s0, 0, R, kp1));
                    ubt = i2(W_AT(s0, 1, R, kc), W_AT(s0, 1, R, kp1));
                } else {
                    ut = i3(W_AT(s0, 0, R, km1), W_AT(s0, 0, R, kc), W_AT(s0, 0, R, kp1), W_AT(s0, 0, R, kp2), sgn(wba + wa));
                    ubt = i3(W_AT(s0, 1, R, km1), W_AT(s0, 1, R, kc), W_AT(s0, 1, R, kp1), W_AT(s0, 1, R, kp2), sgn(wa));
                }
                tf[FU] = wa * ubt + (wba + wa) * ut;
            }
            {  // v_top
                T wa = T(0.5) * (W_AT(s0, 4, R, kp1) + W_AT(s0, 4, R - 1, kp1));
                T wba = T(0.5) * (W_AT(s0, 5, R, kp1) + W_AT(s0, 5, R - 1, kp1));
                T vt, vbt;
                if (low) {
                    vt = i2(W_AT(s0, 2, R, kc), W_AT(s0, 2, R, kp1));
                    vbt = i2(W_AT(s0, 3, R, kc), W_AT(s0, 3, R, kp1));
                } else {
                    vt = i3(W_AT(s0, 2, R, km1), W_AT(s0, 2, R, kc), W_AT(s0, 2, R, kp1), W_AT(s0, 2, R, kp2), sgn(wba + wa));
                    vbt = i3(W_AT(s0, 3, R, km1), W_AT(s0, 3, R, kc), W_AT(s0, 3, R, kp1), W_AT(s0, 3, R, kp2), sgn(wa));
                }
                tf[FV] = wa * vbt + (wba + wa) * vt;
            }
            T W1 = W_AT(s0, 4, R, kp1), WB1 = W_AT(s0, 5, R, kp1);
            auto s_top = [&](int fs, int fb) -> T {
                T st, bt;
                if (low) {
                    st = i2(S_AT(fs, R, kc), S_AT(fs, R, kp1));
                    bt = i2(S_AT(fb, R, kc), S_AT(fb, R, kp1));
                } else {
                    st = i3(S_AT(fs, R, km1), S_AT(fs, R, kc), S_AT(fs, R, kp1), S_AT(fs, R, kp2), sgn(WB1 + W1));
                    bt = i3(S_AT(fb, R, km1), S_AT(fb, R, kc), S_AT(fb, R, kp1), S_AT(fb, R, kp2), sgn(W1));
                }
                return W1 * bt + (WB1 + W1) * st;
            };
            tf[FTH] = s_top(0, 1);
            if (MOIST) {
                tf[FQV] = s_top(2, 3);
                T wb = WB1 + W1;
                T t = low ? i2(S_AT(4, R, kc), S_AT(4, R, kp1)) : i3(S_AT(4, R, km1), S_AT(4, R, kc), S_AT(4, R, kp1), S_AT(4, R, kp2), sgn(wb));
                tf[FQC] = wb * t;
            }
            if (k == 0) {
#pragma unroll
                for (int f = 0; f < NF; f++) tf[f] = T(0);
            }
        }
#pragma unroll
        for (int f = 0; f < NF; f++) tface[(f * TJ + r) * KP + k] = tf[f];
        __syncthreads();
        // rain top faces need both levels' fall speeds
        T qrt = T(0), qrb = T(0);
        if (MOIST) {
            const int kp1 = min(kc + 1, NZ - 1), km1 = max(kc - 1, 0), kp2 = min(kc + 2, NZ - 1);
            auto vt_at = [&](int kk) -> T { return kk == 0 ? vtp[r * KP + 1] : vtp[r * KP + kk]; };
            auto qr_top = [&](int kk) -> T {
                // top face of level kk of this column (fluxes.cpp:734, 780; 966-1030)
                if (kk == 0) {
                    T t0 = T(0.5) * (S_AT(5, R, 0) + S_AT(5, R, 1));
                    return -T(0.5) * (vt_at(1) + vt_at(0)) * t0;
                }
                int a1 = min(kk + 1, NZ - 1), am1 = max(kk - 1, 0), a2 = min(kk + 2, NZ - 1);
                T vk = vt_at(kk), vk1 = vt_at(a1);
                T wb = W_AT(s0, 5, R, a1) + W_AT(s0, 4, R, a1);
                T t;
                if (kk == 1 || kk == NZ - 2) t = i2(S_AT(5, R, kk), S_AT(5, R, a1));
                else t = i3(S_AT(5, R, am1), S_AT(5, R, kk), S_AT(5, R, a1), S_AT(5, R, a2), sgn(W_AT(s0, 5, R, a1) + W_AT(s0, 4, R, a1) - T(0.5) * (vk + vk1)));
                return (wb - T(0.5) * (vk1 + vk)) * t;
            };
            if (active) {
                qrt = qr_top(kc);
                qrb = qr_top(km1);
            }
            (void)kp1;
            (void)kp2;
        }
        // ---- tendencies ---------------------------------------------------------------------
        if (active) {
            const long long p = gidx(i, j, k);
            auto SOUTH = [&](int f) { return nface[(f * (TJ + 1) + r) * KP + k]; };
            auto BOTTOM = [&](int f) { return tface[(f * TJ + r) * KP + k - 1]; };
            {  // u
                T other = T(0.25) * (W_AT(s0, 2, R, k) + W_AT(s0, 2, R + 1, k) + W_AT(sm1, 2, R, k) + W_AT(sm1, 2, R + 1, k));
                T tend = -dtx * (east[FU] - west[FU]) - dty * (nf[FU] - SOUTH(FU)) - DTZ * (rw1 * tf[FU] - rw0 * BOTTOM(FU)) * odr +
                         dt * fj * other;
                ((T *)a.out[0])[p] = ((const T *)a.mprev[0])[p] + step * tend;
            }
            {  // v
                const int sp = slot_of(i + 1);
                T other = T(0.25) * (W_AT(s0, 0, R, k) + W_AT(sp, 0, R, k) + W_AT(s0, 0, R - 1, k) + W_AT(sp, 0, R - 1, k));
                T tend = -dtx * (east[FV] - west[FV]) - dty * (nf[FV] - SOUTH(FV)) - DTZ * (rw1 * tf[FV] - rw0 * BOTTOM(FV)) * odr -
                         dt * fj * other;
                ((T *)a.out[1])[p] = ((const T *)a.mprev[1])[p] + step * tend;
            }
            if (k >= 2) {  // w: buoyancy (advection.cpp:329-334)
                T bk, bkm;
                if (!MOIST) {
                    bk = S_AT(0, R, k) / T(c.tbv[k]);
                    bkm = S_AT(0, R, k - 1) / T(c.tbv[k - 1]);
                } else {
                    bk = S_AT(0, R, k) / T(c.tbv[k]) + T(0.61) * S_AT(2, R, k) - S_AT(4, R, k) - S_AT(5, R, k);
                    bkm = S_AT(0, R, k - 1) / T(c.tbv[k - 1]) + T(0.61) * S_AT(2, R, k - 1) - S_AT(4, R, k - 1) - S_AT(5, R, k - 1);
                }
                ((T *)a.out[2])[p] = ((const T *)a.mprev[2])[p] + step * dt * T(kGrav) * T(0.5) * (bk + bkm);
            }
            // Z_ADVECT_BASE (advection.cpp:16-18)
            auto zbase = [&](const double *base) -> T {
                return (T(0.50) * (T(c.rhow[k + 1]) * W_AT(s0, 4, R, k + 1) * (T(base[k + 1]) - T(base[k])) +
                                   T(c.rhow[k]) * W_AT(s0, 4, R, k) * (T(base[k]) - T(base[k - 1]))) *
                        T(c.one_d_rhou[k]));
            };
            {  // theta
                T tend = -dtx * (east[FTH] - west[FTH]) - dty * (nf[FTH] - SOUTH(FTH)) -
                         DTZ * (rw1 * tf[FTH] - rw0 * BOTTOM(FTH)) * odr - DTZ * zbase(c.tb);
                ((T *)a.out[3])[p] = ((const T *)a.mprev[3])[p] + step * tend;
            }
            if (MOIST) {
                T tq = -dtx * (east[FQV] - west[FQV]) - dty * (nf[FQV] - SOUTH(FQV)) - DTZ * (rw1 * tf[FQV] - rw0 * BOTTOM(FQV)) * odr -
                       DTZ * zbase(c.qb);
                T qv = ((const T *)a.mprev[4])[p] + step * tq;
                T tc = -dtx * (east[FQC] - west[FQC]) - dty * (nf[FQC] - SOUTH(FQC)) - DTZ * (rw1 * tf[FQC] - rw0 * BOTTOM(FQC)) * odr;
                T qc = ((const T *)a.mprev[5])[p] + step * tc;
                T tr = -dtx * (east[FQR] - west[FQR]) - dty * (nf[FQR] - SOUTH(FQR)) - DTZ * (rw1 * qrt - rw0 * qrb) * odr;
                T qrv = ((const T *)a.mprev[6])[p] + step * tr;
                // zero_moisture for the advected points (microphysics.cpp:677-698)
                if (qc < T(0)) qc = T(0);
                if (qrv < T(0)) qrv = T(0);
                T qbp = QB[p];
                if (qv + qbp + T(c.qb[k]) < T(0)) qv = -(qbp + T(c.qb[k]));
                ((T *)a.out[4])[p] = qv;
                ((T *)a.out[5])[p] = qc;
                ((T *)a.out[6])[p] = qrv;
                T *VTo = (T *)a.vt;
                VTo[p] = vtk;
                if (k == 1) VTo[p - 1] = vtk;
            }
        }
        // ---- advance the march ----------------------------------------------------------------
#pragma unroll
        for (int f = 0; f < NF; f++) west[f] = east[f];
        __syncthreads();  // all reads of plane i-1 and of the scalar plane are done
        if (i + 1 < i1) {
#pragma unroll
            for (int f = 0; f < nq; f++) {
#pragma unroll
                for (int d = 0; d < 6; d++) q[f][d] = q[f][d + 1];
                q[f][6] = qsrc[f][gidx(i + 4, jc, kc)];
            }
            load_win_plane(i + 2);
        }
    }
}

// zero_moisture outside the advected box (the fused kernel clips the points it writes)
template <typename T>
__global__ void k_zero_moisture_shell(Geo g, Cols c, T *__restrict__ QVP, T *__restrict__ QCP, T *__restrict__ QRP,
                                      const T *__restrict__ QB, long long n, int bil, int bih, int bjl, int bjh, int il, int ih, int jl,
                                      int jh) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int k = (int)(p % g.NZ);
    long long rr = p / g.NZ;
    int j = (int)(rr % g.NY), i = (int)(rr / g.NY);
    bool inbox = i >= bil && i < bih && j >= bjl && j < bjh && k >= 1 && k < g.NZ - 1;
    if (inbox) return;
    if (QCP[p] < T(0)) QCP[p] = T(0);
    if (QRP[p] < T(0)) QRP[p] = T(0);
    if (i >= il && i < ih && j >= jl && j < jh && k >= 1 && k < g.NZ - 1) {
        if (QVP[p] + QB[p] + T(c.qb[k]) < T(0)) QVP[p] = -(QB[p] + T(c.qb[k]));
    }
}

template <typename T, int TJ>
int stage_t(Ctx *c, double step, int il, int ih, int jl, int jh) {
    const bool moist = c->g.microphysics != 0;
    FusedArgs a{};
    for (int f = 0; f < 7; f++) {
        a.cur[f] = c->prog[f][DYN_ROLE_CUR];
        a.mprev[f] = c->prog[f][DYN_ROLE_M];
        a.out[f] = c->prog[f][DYN_ROLE_P];
    }
    for (int f = 0; f < 5; f++) a.bar[f] = c->bar[f];
    a.vt = c->vt;
    a.il = il;
    a.ih = ih;
    a.jl = jl;
    a.jh = jh;
    a.ti = 16;
    dim3 blk(KP, TJ), grd((jh - jl + TJ - 1) / TJ, (ih - il + a.ti - 1) / a.ti);
    if (moist) {
        size_t sh = Smem<T, TJ, true>::bytes;
        Timed _t(c, "fused_advect", (double)(ih - il) * (jh - jl) * (c->geo.NZ - 2));
        hipLaunchKernelGGL((k_stage_advect<T, TJ, true>), grd, blk, sh, c->stream, c->geo, c->cols, a, T(step));
    } else {
        size_t sh = Smem<T, TJ, false>::bytes;
        Timed _t(c, "fused_advect", (double)(ih - il) * (jh - jl) * (c->geo.NZ - 2));
        hipLaunchKernelGGL((k_stage_advect<T, TJ, false>), grd, blk, sh, c->stream, c->geo, c->cols, a, T(step));
    }
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T>
int shell_t(Ctx *c, int il, int ih, int jl, int jh) {
    long long n = (long long)c->npts;
    Timed _t(c, "zero_moisture_shell", (double)n);
    hipLaunchKernelGGL(k_zero_moisture_shell<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->geo, c->cols,
                       (T *)c->prog[DYN_QV][2], (T *)c->prog[DYN_QC][2], (T *)c->prog[DYN_QR][2], (const T *)c->bar[4], n, 3,
                       c->g.nx - 3, 3, c->g.ny - 3, il, ih, jl, jh);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}


// ---- boundary epilogues of a stage (PGF + apply_boundary_condition(_microphysics)) ------------------
// A wavefront owns one source column isrc in [3, NX-3) (lane = k) and writes that column's final values
// and, for the six columns the periodic copy feeds (boundaries.cpp:562-574), the halo column's values.
// Per field the reference's order is kept: sponge rows (boundaries.cpp:466-475), periodic copy, then the
// lid from i = 1, j = 1 (boundaries.cpp:304-326); column 0 therefore keeps the pre-lid values.  Lid
// values come from lanes 1 and NZ-2 through cross-lane reads, so no wavefront reads another's output.
struct PostArgs {
    void *f[7];          // UP VP WP THP QVP QCP QRP (forecast buffers)
    const void *pi, *qb;  // PI, QBAR
    const double *cs, *cn;
    int js, jn;
    int zl, zh, zjl, zjh;  // zero_moisture bounds (qv clip)
};

template <typename T>
__device__ __forceinline__ T sponge_coef(const Geo &g, const PostArgs &a, int j) {
    if (j < a.js) return T(a.cs[j]);
    if (j >= g.NY - a.jn) return T(a.cn[j - (g.NY - a.jn)]);
    return T(-1);  // no sponge
}

template <typename T>
__device__ __forceinline__ void write_col(const Geo &g, T *A, int isrc, int j, int k, T fin, T prelid) {
    long long s = g.NYNZ;
    long long p = (long long)isrc * s + (long long)j * g.NZ + k;
    A[p] = fin;
    const int shift = g.NX - 6;
    if (isrc >= shift && isrc < shift + 3) {
        int ih = isrc - shift;  // halo 0, 1, 2
        A[p - (long long)shift * s] = ih == 0 ? prelid : fin;
    }
    if (isrc >= 3 && isrc < 6) A[p + (long long)shift * s] = fin;
}

// PGF (solver.cpp:123-135) + sponge/periodic/lid for UP, VP and WP
template <typename T>
__global__ void __launch_bounds__(256) k_post_uvw(Geo g, Cols c, PostArgs a, T step, bool pgf) {
    const int k = threadIdx.x;
    const int j = blockIdx.x * blockDim.y + threadIdx.y;
    const int isrc = blockIdx.y + 3;
    if (j >= g.NY) return;
    const int NZ = g.NZ;
    const int kc = k < NZ ? k : NZ - 1;
    const long long p = (long long)isrc * g.NYNZ + (long long)j * NZ + kc;
    T *UP = (T *)a.f[0], *VP = (T *)a.f[1], *WP = (T *)a.f[2];
    const T *PI = (const T *)a.pi;
    T u = UP[p], v = VP[p], w = WP[p];
    if (pgf && j >= 1 && j <= g.NY - 2 && kc >= 1 && kc <= NZ - 2) {
        u -= step * (T(g.dtx) * (PI[p] - PI[p - g.NYNZ]));
        v -= step * (T(g.dty) * (PI[p] - PI[p - NZ]));
        if (kc >= 2) w -= step * ((T(g.dtz) * T(c.mw[kc])) * (PI[p] - PI[p - 1]));
    }
    T coef = sponge_coef<T>(g, a, j);
    if (coef >= T(0) && kc >= 1) {
        u -= coef * u;
        v -= coef * v;
        w -= coef * w;
    }
    // lid from lanes 1 and NZ-2 (j >= 1; every source column has i >= 1)
    T u1 = __shfl(u, 1), uz = __shfl(u, NZ - 2), v1 = __shfl(v, 1), vz = __shfl(v, NZ - 2);
    if (k >= NZ) return;
    T uf = u, vf = v, wf = w;
    if (j >= 1) {
        if (k == 0) { uf = u1; vf = v1; }
        if (k == NZ - 1) { uf = uz; vf = vz; }
        if (k == 0 || k == 1 || k == NZ - 1) wf = T(0);
    }
    write_col(g, UP, isrc, j, k, uf, u);
    write_col(g, VP, isrc, j, k, vf, v);
    write_col(g, WP, isrc, j, k, wf, w);
}

// sponge/periodic/lid for THP (and QVP, QCP, QRP with zero_moisture outside the advected box); only the
// columns a sponge row, a ghost row or the periodic copy touches are rewritten whole, the others get the
// two lid levels
template <typename T, bool MOIST>
__global__ void __launch_bounds__(256) k_post_scalars(Geo g, Cols c, PostArgs a, bool clip) {
    const int k = threadIdx.x;
    const int j = blockIdx.x * blockDim.y + threadIdx.y;
    const int isrc = blockIdx.y + 3;
    if (j >= g.NY) return;
    const int NZ = g.NZ;
    const int kc = k < NZ ? k : NZ - 1;
    const long long col = (long long)isrc * g.NYNZ + (long long)j * NZ;
    const T coef = sponge_coef<T>(g, a, j);
    const bool full = coef >= T(0) || j < 3 || j >= g.NY - 3 || isrc < 6 || isrc >= g.NX - 6;
    constexpr int NFLD = MOIST ? 4 : 1;
    if (!full) {  // interior column: lid only (values at 1 and NZ-2 are final already)
        if (j >= 1 && (k == 0 || k == NZ - 1)) {
            int ks = k == 0 ? 1 : NZ - 2;
#pragma unroll
            for (int f = 0; f < NFLD; f++) {
                T *A = (T *)a.f[3 + f];
                A[col + k] = A[col + ks];
            }
        }
        return;
    }
#pragma unroll
    for (int f = 0; f < NFLD; f++) {
        T *A = (T *)a.f[3 + f];
        T x = A[col + kc];
        if (MOIST && clip && f == 2 && x < T(0)) x = T(0);  // qc
        if (MOIST && clip && f == 3 && x < T(0)) x = T(0);  // qr
        if (MOIST && clip && f == 1 && isrc >= a.zl && isrc < a.zh && j >= a.zjl && j < a.zjh && kc >= 1 && kc <= NZ - 2) {
            T qbp = ((const T *)a.qb)[col + kc];
            if (x + qbp + T(c.qb[kc]) < T(0)) x = -(qbp + T(c.qb[kc]));
        }
        if (coef >= T(0) && kc >= 1) x -= coef * x;
        T x1 = __shfl(x, 1), xz = __shfl(x, NZ - 2);
        if (k < NZ) {
            T fin = x;
            if (j >= 1) {
                if (k == 0) fin = x1;
                if (k == NZ - 1) fin = xz;
            }
            write_col(g, A, isrc, j, k, fin, x);
        }
    }
}

template <typename T>
int post_t(Ctx *c, double step, bool pgf, bool scalars, bool clip) {
    PostArgs a{};
    for (int f = 0; f < 7; f++) a.f[f] = c->prog[f][DYN_ROLE_P];
    a.pi = c->pi;
    a.qb = c->bar[4];
    a.cs = c->sponge_s;
    a.cn = c->sponge_n;
    a.js = c->g.sponge_south;
    a.jn = c->g.sponge_north;
    a.zl = 1;
    a.zh = c->g.nx - 1;
    a.zjl = 1;
    a.zjh = c->g.ny - 1;
    const int rows = 4;
    dim3 blk(KP, rows), grd((c->g.ny + rows - 1) / rows, c->g.nx - 6);
    {
        Timed _t(c, "post_uvw", (double)c->npts);
        hipLaunchKernelGGL(k_post_uvw<T>, grd, blk, 0, c->stream, c->geo, c->cols, a, T(step), pgf);
    }
    DYN_LAUNCH_CHECK();
    if (scalars) {
        Timed _t(c, "post_scalars", (double)c->npts);
        if (c->g.microphysics) hipLaunchKernelGGL((k_post_scalars<T, true>), grd, blk, 0, c->stream, c->geo, c->cols, a, clip);
        else hipLaunchKernelGGL((k_post_scalars<T, false>), grd, blk, 0, c->stream, c->geo, c->cols, a, clip);
        DYN_LAUNCH_CHECK();
    }
    return DYN_OK;
}

}  // namespace

bool fused_supported(const Ctx *c) { return c->g.nz <= KP; }

// the stage kernel uses more than 64 KiB of LDS: raise the per-launch limit once, outside any capture
int fused_prepare(Ctx *c) {
    (void)c;
    DYN_HIP(hipFuncSetAttribute((const void *)k_stage_advect<float, 8, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)Smem<float, 8, true>::bytes));
    DYN_HIP(hipFuncSetAttribute((const void *)k_stage_advect<float, 8, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)Smem<float, 8, false>::bytes));
    DYN_HIP(hipFuncSetAttribute((const void *)k_stage_advect<double, 4, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)Smem<double, 4, true>::bytes));
    DYN_HIP(hipFuncSetAttribute((const void *)k_stage_advect<double, 4, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)Smem<double, 4, false>::bytes));
    return DYN_OK;
}

// advect_uvw_velocity + advect_theta (+ advect_microphysics_cell and the interior zero_moisture)
// over the box [il,ih) x [jl,jh)
int launch_fused_advect(Ctx *c, double step, int il, int ih, int jl, int jh) {
    if (c->prec == DYN_FP32) return stage_t<float, 8>(c, step, il, ih, jl, jh);
    return stage_t<double, 4>(c, step, il, ih, jl, jh);
}

// PGF (optional) and the boundary update of every prognostic after launch_fused_advect; stands for
// dyn_pgf + zero_moisture's remainder + dyn_apply_bc + dyn_apply_bc_micro of one stage
int launch_post_stage(Ctx *c, double step, bool pgf, bool scalars, bool clip) {
    if (c->prec == DYN_FP32) return post_t<float>(c, step, pgf, scalars, clip);
    return post_t<double>(c, step, pgf, scalars, clip);
}

// the rest of zero_moisture(il,ih,jl,jh) after launch_fused_advect on the box [3,NX-3) x [3,NY-3)
int launch_zero_moisture_shell(Ctx *c, int il, int ih, int jl, int jh) {
    if (c->prec == DYN_FP32) return shell_t<float>(c, il, ih, jl, jh);
    return shell_t<double>(c, il, ih, jl, jh);
}

}  // namespace dyn
