// Fused stage advection for gfx950: advect_uvw_velocity + advect_theta + advect_microphysics_cell
// (+ the fall speed and the interior part of zero_moisture) in one pass over HBM.
//
// Mapping: one 64-lane wavefront per j-row, lane = level k (NZ <= 64; two wavefronts for 66 <= NZ <= 128,
// which hand the values crossing levels 63/64 over in LDS); a block of TJ rows marches
// along i over a chunk of TI columns, as the reference's slab loop does (advection.cpp:313-325):
//   * i-stencils come from per-thread register queues (own j,k; columns i-3..i+3) and the east-face
//     flux of column i is carried as the west face of column i+1 (fluxes.cpp:891,1322,1504);
//   * j- and k-stencils read LDS planes of TJ+6 rows; U/UB are kept for planes i and i+1, V/VB/W/WB for
//     planes i-1 and i (staggered advecting velocities), the scalars for plane i;
//   * north and top faces are computed once per point and exchanged through LDS as the south and
//     bottom faces of the neighbours.
// Every face and tendency uses the same expression tree as the per-operator kernels, so the fused and
// unfused paths agree to the last bit.
#include <algorithm>

#include "dyn_internal.hpp"

namespace dyn {
namespace {

constexpr int KP = 64;  // lanes per row (levels)

template <typename T>
__device__ __forceinline__ T sgn(T x) { return x > T(0) ? T(1) : T(-1); }

// interpolants of fluxes.cpp:110-206 (face_i5, face_i3)
template <typename T>
__device__ __forceinline__ T i5(T pm2, T pm1, T p0, T p1, T p2, T p3, T sign) { return face_i5(pm2, pm1, p0, p1, p2, p3, sign); }
template <typename T>
__device__ __forceinline__ T i3(T pm1, T p0, T p1, T p2, T sign) { return face_i3(pm1, p0, p1, p2, sign); }
template <typename T>
__device__ __forceinline__ T i2(T p0, T p1) { return T(0.5) * (p1 + p0); }

// Neighbouring levels of a column sit in neighbouring lanes.  Level k+1 and level k-1 come by the
// wavefront shifts of DPP (wave_shl:1 and wave_shr:1, a modifier of a VALU move: no LDS round trip); a
// lane whose source lies outside the wavefront keeps its own value.  That is the clamp the callers want at
// level 0 and at the last lane, and the wavefront boundary of a two-wavefront column is overwritten from
// LDS by the callers.  Lanes past NZ-1 hold the values loaded at level NZ-1 (clamped loads) or values no
// active lane reads.  A fixed level comes by readlane (lane index uniform).
__device__ __forceinline__ int dpp_next(int x) { return __builtin_amdgcn_update_dpp(x, x, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_prev(int x) { return __builtin_amdgcn_update_dpp(x, x, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_next(float x) { return __int_as_float(dpp_next(__float_as_int(x))); }
__device__ __forceinline__ float lane_prev(float x) { return __int_as_float(dpp_prev(__float_as_int(x))); }
__device__ __forceinline__ float lane_read(float x, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane)); }
__device__ __forceinline__ double lane_next(double x) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)dpp_next((int)b), hi = (unsigned)dpp_next((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double lane_prev(double x) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)dpp_prev((int)b), hi = (unsigned)dpp_prev((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double lane_read(double x, int lane) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, lane), hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Element offsets are 32-bit (the host checks that a field is below 4 GiB): the base pointer stays in
// SGPRs and one VGPR byte offset serves every field at the same point.
template <typename T>
__device__ __forceinline__ T ld(const void *base, unsigned idx) {
    return *(const T *)((const char *)base + (size_t)(idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st(void *base, unsigned idx, T v) {
    *(T *)((char *)base + (size_t)(idx * (unsigned)sizeof(T))) = v;
}

// Tile t = jblock + nbj * ichunk, placed by xcd_tile: neighbouring j-blocks of one i-chunk (which share
// halo rows) sit behind the same L2.

struct StageArgs {
    const void *cur[7];   // U V W TH QV QC QR (current)
    const void *mprev[7]; // M buffers
    void *out[7];         // P buffers
    const void *bar[5];   // UB VB WB THB QB
    void *vt;
    int il, ih, jl, jh;   // advected box
    int ti;               // columns one block marches
    int nbj;              // j-blocks
};

// ---- momentum: advect_uvw_velocity (u, v fluxes and Coriolis; w = buoyancy, advection.cpp:304-339) -------
// LDS: U UB (planes i, i+1) and V VB (planes i-1, i) on rows j0-3 .. j0+TJ+2, W WB (planes i-1, i) on rows
// j0-1 .. j0+TJ-1, each in a ring of three slots (plane p in slot p % 3); north faces of rows j0-1 .. j0+TJ-1.
// Register queues hold U UB V VB of the own row at columns i-2 .. i+3.
template <typename T, int TJ, int KW = 1>
struct MomLds {
    static constexpr int KPW = KP * KW;
    static constexpr int R = TJ + 6;
    static constexpr int RW = TJ + 1;
    static constexpr size_t uv = 3 * 4 * (size_t)R * KPW;    // [slot][U UB V VB][row][k]
    static constexpr size_t w = 3 * 2 * (size_t)RW * KPW;    // [slot][W WB][row][k]
    static constexpr size_t nf = 2 * (size_t)(TJ + 1) * KPW; // [u v][row + 1][k]
    static constexpr size_t x = KW == 2 ? 3 * (size_t)TJ : 0;  // [row][bu bv bk] from level 63 to 64
    static constexpr size_t bytes = (uv + w + nf + x) * sizeof(T);
};

// one tile (TJ rows x ti columns x all levels) of the momentum stage; every thread of the block calls it
template <typename T, int TJ, bool MOIST, int KW = 1>
__device__ __forceinline__ void adv_momentum_tile(const Geo &g, const Cols &c, const StageArgs &a, T step, int t, unsigned char *smem_raw) {
    using L = MomLds<T, TJ, KW>;
    constexpr int KPW = L::KPW;
    constexpr int R = L::R, RW = L::RW;
    constexpr int HH = (6 + TJ - 1) / TJ, HW = (RW + TJ - 1) / TJ;  // halo / W plane rows per wavefront
    T *uvp = (T *)smem_raw;
    T *wpl = uvp + L::uv;
    T *nface = wpl + L::w;
    T *xch = nface + L::nf;

    const int k = threadIdx.x, r = threadIdx.y;
    const int j0 = a.jl + (t % a.nbj) * TJ;
    const int i0 = a.il + (t / a.nbj) * a.ti;
    const int i1 = min(i0 + a.ti, a.ih);
    const int NZ = g.NZ, NY = g.NY;
    const unsigned NYNZ = (unsigned)g.NYNZ;
    const int j = j0 + r;
    const bool active = j < a.jh && k >= 1 && k <= NZ - 2;
    const int kc = min(k, NZ - 1), jc = min(j, NY - 1);
    const int km1 = max(kc - 1, 0), kp1 = min(kc + 1, NZ - 1), kp2 = min(kc + 2, NZ - 1);
    const unsigned rowoff = (unsigned)(jc * NZ + kc);
    const int RO = r + 3, WO = r + 1;  // own row in the U/V and W planes

    auto UV = [&](int slot, int f, int row) -> T * { return uvp + ((slot * 4 + f) * R + row) * KPW; };
    auto WW = [&](int slot, int f, int row) -> T * { return wpl + ((slot * 2 + f) * RW + row) * KPW; };
    const void *qsrc[4] = {a.cur[0], a.bar[0], a.cur[1], a.bar[1]};
    T q[4][6], qs[4];  // U UB V VB of the own row at columns i-2 .. i+3 (i-3 .. i+2 in the prologue)

    // plane rows: U UB of plane iu, V VB W WB of plane iv.  The own row of U UB V VB comes from the
    // queues (which hold those columns already); only the six halo rows j0-3..j0-1, j0+TJ..j0+TJ+2 are
    // loaded
    auto halo_row = [&](int hh) { return hh < 3 ? hh : TJ + hh; };
    T su[HH][4], sw[HW][2];
    auto fetch_planes = [&](int iu, int iv) {
#pragma unroll
        for (int h = 0; h < HH; h++) {
            const int hh = r + h * TJ;
            if (hh < 6) {
                const int rr = halo_row(hh);
                const int jj = min(max(j0 - 3 + rr, 0), NY - 1);
                const unsigned pu = (unsigned)iu * NYNZ + (unsigned)(jj * NZ + kc), pv = (unsigned)iv * NYNZ + (unsigned)(jj * NZ + kc);
                su[h][0] = ld<T>(a.cur[0], pu);
                su[h][1] = ld<T>(a.bar[0], pu);
                su[h][2] = ld<T>(a.cur[1], pv);
                su[h][3] = ld<T>(a.bar[1], pv);
            }
        }
#pragma unroll
        for (int h = 0; h < HW; h++) {
            const int rr = r + h * TJ;
            if (rr < RW) {
                const int jj = min(max(j0 - 1 + rr, 0), NY - 1);
                const unsigned pv = (unsigned)iv * NYNZ + (unsigned)(jj * NZ + kc);
                sw[h][0] = ld<T>(a.cur[2], pv);
                sw[h][1] = ld<T>(a.bar[2], pv);
            }
        }
    };
    // qu, qv: queue positions of columns iu and iv
    auto store_planes = [&](int iu, int iv, int qu, int qv) {
        const int su_slot = iu % 3, sv_slot = iv % 3;
        UV(su_slot, 0, RO)[k] = q[0][qu];
        UV(su_slot, 1, RO)[k] = q[1][qu];
        UV(sv_slot, 2, RO)[k] = q[2][qv];
        UV(sv_slot, 3, RO)[k] = q[3][qv];
#pragma unroll
        for (int h = 0; h < HH; h++) {
            const int hh = r + h * TJ;
            if (hh < 6) {
                const int rr = halo_row(hh);
                UV(su_slot, 0, rr)[k] = su[h][0];
                UV(su_slot, 1, rr)[k] = su[h][1];
                UV(sv_slot, 2, rr)[k] = su[h][2];
                UV(sv_slot, 3, rr)[k] = su[h][3];
            }
        }
#pragma unroll
        for (int h = 0; h < HW; h++) {
            const int rr = r + h * TJ;
            if (rr < RW) {
                WW(sv_slot, 0, rr)[k] = sw[h][0];
                WW(sv_slot, 1, rr)[k] = sw[h][1];
            }
        }
    };
    // own-column inputs of column i: the buoyancy term's theta', qv, qc, qr and the M values of u, v, w
    T cb[7] = {}, sb[7] = {};
    auto fetch_col = [&](int i, T *b) {
        const unsigned p = (unsigned)i * NYNZ + rowoff;
        b[0] = ld<T>(a.cur[3], p);
        if (MOIST) {
            b[1] = ld<T>(a.cur[4], p);
            b[2] = ld<T>(a.cur[5], p);
            b[3] = ld<T>(a.cur[6], p);
        }
        b[4] = ld<T>(a.mprev[0], p);
        b[5] = ld<T>(a.mprev[1], p);
        b[6] = ld<T>(a.mprev[2], p);
    };

    // prologue: queues at columns i0-3 .. i0+2, planes U(i0, i0+1), V/W(i0-1, i0)
#pragma unroll
    for (int f = 0; f < 4; f++)
#pragma unroll
        for (int d = 0; d < 6; d++) q[f][d] = ld<T>(qsrc[f], (unsigned)(i0 - 3 + d) * NYNZ + rowoff);
    fetch_planes(i0, i0 - 1);
    store_planes(i0, i0 - 1, 3, 2);
    fetch_planes(i0 + 1, i0);
    store_planes(i0 + 1, i0, 4, 3);
    fetch_col(i0, cb);

    const T dtx = T(g.dtx), dty = T(g.dty), dt = T(g.dt);
    const T DTZ = T(g.dtz) * T(c.mu[kc]);
    const T rw1 = T(c.rhow[kp1]), rw0 = T(c.rhow[kc]), odr = T(c.one_d_rhou[kc]);
    const T fj = T(c.f[jc]);
    // read once: a load inside the march would be waited for together with the next column's prefetch
    const T tbvk = T(c.tbv[kc]);
    const bool low = (k == 1 || k == NZ - 2);

    // east faces of the queue's centre column (u_east, v_east; fluxes.cpp:1384-1592); Ujm1/UBjm1: row j-1 of
    // the U planes at the column east of the centre
    auto east_faces = [&](const T *Ujm1, const T *UBjm1, T *e) {
        {
            T ua = T(0.5) * (q[0][3] + q[0][2]);
            T uba = T(0.5) * (q[1][3] + q[1][2]);
            T ue = i5(q[0][0], q[0][1], q[0][2], q[0][3], q[0][4], q[0][5], sgn(uba + ua));
            T ube = i5(q[1][0], q[1][1], q[1][2], q[1][3], q[1][4], q[1][5], sgn(ua));
            e[0] = ua * ube + (uba + ua) * ue;
        }
        {
            T ua = T(0.5) * (q[0][3] + Ujm1[k]);
            T uba = T(0.5) * (q[1][3] + UBjm1[k]);
            T ve = i5(q[2][0], q[2][1], q[2][2], q[2][3], q[2][4], q[2][5], sgn(uba + ua));
            T vbe = i5(q[3][0], q[3][1], q[3][2], q[3][3], q[3][4], q[3][5], sgn(ua));
            e[1] = ua * vbe + (uba + ua) * ve;
        }
    };
    __syncthreads();
    T west[2], east[2];
    east_faces(UV(i0 % 3, 0, RO - 1), UV(i0 % 3, 1, RO - 1), west);  // column i0-1
#pragma unroll
    for (int f = 0; f < 4; f++) {
#pragma unroll
        for (int d = 0; d < 5; d++) q[f][d] = q[f][d + 1];
        q[f][5] = ld<T>(qsrc[f], (unsigned)(i0 + 3) * NYNZ + rowoff);
    }
    // Every prologue load has landed before the march starts: a load still outstanding at the loop's
    // entry makes the compiler wait, inside every iteration, for the next column's prefetch as well.
    __builtin_amdgcn_s_waitcnt(0);

    for (int i = i0; i < i1; i++) {
        const bool more = i + 1 < i1;
        const int c0 = i % 3, cm1 = (i + 2) % 3, cp1 = (i + 1) % 3;
        const unsigned p = (unsigned)i * NYNZ + rowoff;
        // loads for the next column go out first, its M values included
        if (more) {
#pragma unroll
            for (int f = 0; f < 4; f++) qs[f] = ld<T>(qsrc[f], (unsigned)(i + 4) * NYNZ + rowoff);
            fetch_planes(i + 2, i + 1);
            fetch_col(i + 1, sb);
        }
        const T um = cb[4], vm = cb[5], wm = cb[6];

        east_faces(UV(cp1, 0, RO - 1), UV(cp1, 1, RO - 1), east);
        // north faces (u_north, v_north) of row RR of the U/V planes
        auto north_faces = [&](int RR, T *nf) {
            const T *U0 = UV(c0, 0, 0), *UB0 = UV(c0, 1, 0), *V0 = UV(c0, 2, 0), *VB0 = UV(c0, 3, 0);
            const T *Vm = UV(cm1, 2, 0), *VBm = UV(cm1, 3, 0);
#define PL(ptr, row) ptr[(row) * KPW + k]
            {
                T va = T(0.5) * (PL(V0, RR + 1) + PL(Vm, RR + 1));
                T vba = T(0.5) * (PL(VB0, RR + 1) + PL(VBm, RR + 1));
                T un = i5(PL(U0, RR - 2), PL(U0, RR - 1), PL(U0, RR), PL(U0, RR + 1), PL(U0, RR + 2), PL(U0, RR + 3), sgn(vba + va));
                T ubn = i5(PL(UB0, RR - 2), PL(UB0, RR - 1), PL(UB0, RR), PL(UB0, RR + 1), PL(UB0, RR + 2), PL(UB0, RR + 3), sgn(va));
                nf[0] = va * ubn + (vba + va) * un;
            }
            {
                T va = T(0.5) * (PL(V0, RR + 1) + PL(V0, RR));
                T vba = T(0.5) * (PL(VB0, RR + 1) + PL(VB0, RR));
                T vn = i5(PL(V0, RR - 2), PL(V0, RR - 1), PL(V0, RR), PL(V0, RR + 1), PL(V0, RR + 2), PL(V0, RR + 3), sgn(vba + va));
                T vbn = i5(PL(VB0, RR - 2), PL(VB0, RR - 1), PL(VB0, RR), PL(VB0, RR + 1), PL(VB0, RR + 2), PL(VB0, RR + 3), sgn(va));
                nf[1] = va * vbn + (vba + va) * vn;
            }
#undef PL
        };
        T nf[2];
        north_faces(RO, nf);
        nface[(0 * (TJ + 1) + r + 1) * KPW + k] = nf[0];
        nface[(1 * (TJ + 1) + r + 1) * KPW + k] = nf[1];
        if (r == 0) {
            T sf[2];
            north_faces(RO - 1, sf);
            nface[(0 * (TJ + 1)) * KPW + k] = sf[0];
            nface[(1 * (TJ + 1)) * KPW + k] = sf[1];
        }
        // top faces (u_top, v_top) of this level; level 0 has none
        T tf[2];
        {
            const T *U0 = UV(c0, 0, RO), *UB0 = UV(c0, 1, RO);
            const T *W0 = WW(c0, 0, WO), *WB0 = WW(c0, 1, WO), *Wm = WW(cm1, 0, WO), *WBm = WW(cm1, 1, WO);
            T wa = T(0.5) * (W0[kp1] + Wm[kp1]);
            T wba = T(0.5) * (WB0[kp1] + WBm[kp1]);
            T ut, ubt;
            if (low) {
                ut = i2(U0[kc], U0[kp1]);
                ubt = i2(UB0[kc], UB0[kp1]);
            } else {
                ut = i3(U0[km1], U0[kc], U0[kp1], U0[kp2], sgn(wba + wa));
                ubt = i3(UB0[km1], UB0[kc], UB0[kp1], UB0[kp2], sgn(wa));
            }
            tf[0] = wa * ubt + (wba + wa) * ut;
        }
        {
            const T *V0 = UV(c0, 2, RO), *VB0 = UV(c0, 3, RO);
            const T *W0 = WW(c0, 0, WO), *WB0 = WW(c0, 1, WO), *Wj = WW(c0, 0, WO - 1), *WBj = WW(c0, 1, WO - 1);
            T wa = T(0.5) * (W0[kp1] + Wj[kp1]);
            T wba = T(0.5) * (WB0[kp1] + WBj[kp1]);
            T vt, vbt;
            if (low) {
                vt = i2(V0[kc], V0[kp1]);
                vbt = i2(VB0[kc], VB0[kp1]);
            } else {
                vt = i3(V0[km1], V0[kc], V0[kp1], V0[kp2], sgn(wba + wa));
                vbt = i3(VB0[km1], VB0[kc], VB0[kp1], VB0[kp2], sgn(wa));
            }
            tf[1] = wa * vbt + (wba + wa) * vt;
        }
        if (k == 0) tf[0] = tf[1] = T(0);
        // bottom face of level k = top face of level k-1; buoyancy of level k-1 (advection.cpp:42-58)
        T bu = lane_prev(tf[0]), bv = lane_prev(tf[1]);
        T bk = cb[0] / tbvk;
        if (MOIST) bk = cb[0] / tbvk + T(0.61) * cb[1] - cb[2] - cb[3];
        T bkm = lane_prev(bk);
        // two wavefronts per column: level 64 (the second one's first lane) takes level 63's values from LDS
        if (KW == 2 && k == KP - 1) {
            xch[3 * r] = tf[0];
            xch[3 * r + 1] = tf[1];
            xch[3 * r + 2] = bk;
        }
        __syncthreads();
        if (KW == 2 && k == KP) {
            bu = xch[3 * r];
            bv = xch[3 * r + 1];
            bkm = xch[3 * r + 2];
        }
        T ou = T(0), ov = T(0), ow = T(0);  // this level's new values, stored after the barrier below
        if (active) {
            const T su0 = nface[(0 * (TJ + 1) + r) * KPW + k], sv0 = nface[(1 * (TJ + 1) + r) * KPW + k];
            {  // u (advection.cpp:73-99)
                const T *V0 = UV(c0, 2, 0), *Vm = UV(cm1, 2, 0);
                T other = T(0.25) * (V0[RO * KPW + k] + V0[(RO + 1) * KPW + k] + Vm[RO * KPW + k] + Vm[(RO + 1) * KPW + k]);
                T tend = -dtx * (east[0] - west[0]) - dty * (nf[0] - su0) - DTZ * (rw1 * tf[0] - rw0 * bu) * odr + dt * fj * other;
                ou = um + step * tend;
            }
            {  // v (advection.cpp:104-128)
                const T *U0 = UV(c0, 0, 0), *Up = UV(cp1, 0, 0);
                T other = T(0.25) * (U0[RO * KPW + k] + Up[RO * KPW + k] + U0[(RO - 1) * KPW + k] + Up[(RO - 1) * KPW + k]);
                T tend = -dtx * (east[1] - west[1]) - dty * (nf[1] - sv0) - DTZ * (rw1 * tf[1] - rw0 * bv) * odr - dt * fj * other;
                ov = vm + step * tend;
            }
            ow = wm + step * dt * T(kGrav) * T(0.5) * (bk + bkm);  // advection.cpp:329-337
        }
        west[0] = east[0];
        west[1] = east[1];
        if (more) {
            store_planes(i + 2, i + 1, 4, 3);
#pragma unroll
            for (int f = 0; f < 4; f++) {
#pragma unroll
                for (int d = 0; d < 5; d++) q[f][d] = q[f][d + 1];
                q[f][5] = qs[f];
            }
#pragma unroll
            for (int f = 0; f < 7; f++) cb[f] = sb[f];
        }
        // Drained on every path, so the compiler does not carry this column's prefetch as possibly
        // outstanding into the next iteration (where it would wait on it, and on the stores below, at the
        // first reuse of those registers).  On the main path the plane update has already waited.
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        // stored after the next column's planes are written: a store issued earlier would hold up the
        // wait for the next column's prefetched values (the memory counter completes them in order)
        if (active) {
            st<T>(a.out[0], p, ou);
            st<T>(a.out[1], p, ov);
            if (k >= 2) st<T>(a.out[2], p, ow);
        }
    }
}

// ---- scalars: advect_theta (+ advect_microphysics_cell, the fall speed and the interior zero_moisture) ----
// LDS: theta' thetabar (qv qbar qc qr) of plane i on rows j0-3 .. j0+TJ+2 in a ring of two slots, north
// faces of rows j0-1 .. j0+TJ-1.  Register queues: the same fields of the own row at columns i-2 .. i+3.
// Own-column inputs (u at i+1, v at j+1, w, the M values) are loaded per wavefront.
template <typename T, int TJ, bool MOIST, int KW = 1>
struct ScalLds {
    static constexpr int KPW = KP * KW;
    static constexpr int R = TJ + 6;
    static constexpr int NS = MOIST ? 6 : 2;
    static constexpr int NFS = MOIST ? 4 : 1;
    static constexpr size_t s = 2 * NS * (size_t)R * KPW;          // [slot][field][row][k]
    static constexpr size_t nf = NFS * (size_t)(TJ + 1) * KPW;     // [face][row + 1][k]
    // two wavefronts per column: [row][w wbar] from level 64 to 63, [row][top faces, rain face] from 63 to 64
    static constexpr size_t xw = KW == 2 ? 2 * (size_t)TJ : 0;
    static constexpr size_t xb = KW == 2 ? (NFS + 1) * (size_t)TJ : 0;
    static constexpr size_t bytes = (s + nf + xw + xb) * sizeof(T);
};

template <typename T, int TJ, bool MOIST, int KW = 1>
__device__ __forceinline__ void adv_scalars_tile(const Geo &g, const Cols &c, const StageArgs &a, T step, int t, unsigned char *smem_raw) {
    using L = ScalLds<T, TJ, MOIST, KW>;
    constexpr int KPW = L::KPW;
    constexpr int R = L::R, NS = L::NS, NFS = L::NFS;
    constexpr int HH = (6 + TJ - 1) / TJ;  // halo plane rows per wavefront
    T *spl = (T *)smem_raw;
    T *nface = spl + L::s;
    T *xw = nface + L::nf, *xb = xw + L::xw;

    const int k = threadIdx.x, r = threadIdx.y;
    const int j0 = a.jl + (t % a.nbj) * TJ;
    const int i0 = a.il + (t / a.nbj) * a.ti;
    const int i1 = min(i0 + a.ti, a.ih);
    const int NZ = g.NZ, NY = g.NY;
    const unsigned NYNZ = (unsigned)g.NYNZ;
    const int j = j0 + r;
    const bool active = j < a.jh && k >= 1 && k <= NZ - 2;
    const int kc = min(k, NZ - 1), jc = min(j, NY - 1), jn = min(j + 1, NY - 1);
    const int km1 = max(kc - 1, 0), kp1 = min(kc + 1, NZ - 1), kp2 = min(kc + 2, NZ - 1);
    const unsigned rowoff = (unsigned)(jc * NZ + kc), rowoffn = (unsigned)(jn * NZ + kc);
    const unsigned topoff = (unsigned)(jc * NZ + NZ - 1);
    const int RO = r + 3;

    auto SP = [&](int slot, int f) -> T * { return spl + (slot * NS + f) * R * KPW; };
    const void *src[6] = {a.cur[3], a.bar[3], a.cur[4], a.bar[4], a.cur[5], a.cur[6]};
    T q[NS][6], qs[NS];  // the own row at columns i-2 .. i+3 (i-3 .. i+2 in the prologue)

    // plane rows: the own row comes from the queue, the six halo rows j0-3..j0-1, j0+TJ..j0+TJ+2 are loaded
    auto halo_row = [&](int hh) { return hh < 3 ? hh : TJ + hh; };
    T ss[HH][NS];
    auto fetch_plane = [&](int i) {
#pragma unroll
        for (int h = 0; h < HH; h++) {
            const int hh = r + h * TJ;
            if (hh < 6) {
                const int jj = min(max(j0 - 3 + halo_row(hh), 0), NY - 1);
                const unsigned pp = (unsigned)i * NYNZ + (unsigned)(jj * NZ + kc);
#pragma unroll
                for (int f = 0; f < NS; f++) ss[h][f] = ld<T>(src[f], pp);
            }
        }
    };
    // qd: queue position of column i
    auto store_plane = [&](int i, int qd) {
        const int slot = i % 2;
#pragma unroll
        for (int f = 0; f < NS; f++) SP(slot, f)[RO * KPW + k] = q[f][qd];
#pragma unroll
        for (int h = 0; h < HH; h++) {
            const int hh = r + h * TJ;
            if (hh < 6) {
                const int rr = halo_row(hh);
#pragma unroll
                for (int f = 0; f < NS; f++) SP(slot, f)[rr * KPW + k] = ss[h][f];
            }
        }
    };
    // Register budget: in FP32 the M values of theta and q and qbar travel with the column's prefetch, and
    // a column's results are stored after the end-of-column barrier (so no store is counted ahead of the
    // next column's loads).  In FP64 both would push the kernel further into scratch spills, whose
    // reloads wait on the memory counter too: there the M values are loaded when the column starts and
    // the results stored before the plane update (1.90 against 1.99 ms per moist 256x256x64 step).
    constexpr bool kLean = sizeof(T) == 8;
    // own-column inputs of column i: u, ubar at i+1; v, vbar at j+1 and j; w, wbar; VT at the top level;
    // the M values of theta and q and qbar (FP32)
    struct Col {
        T u1, ub1, vn, vbn, vs, vbs, w0, wb0, vtt, mth, mqv, mqc, mqr, qbp;
    };
    auto fetch_col = [&](int i, Col &o) {
        const unsigned p = (unsigned)i * NYNZ + rowoff, pe = (unsigned)(i + 1) * NYNZ + rowoff, pn = (unsigned)i * NYNZ + rowoffn;
        o.u1 = ld<T>(a.cur[0], pe);
        o.ub1 = ld<T>(a.bar[0], pe);
        o.vn = ld<T>(a.cur[1], pn);
        o.vbn = ld<T>(a.bar[1], pn);
        if (r == 0) {
            o.vs = ld<T>(a.cur[1], p);
            o.vbs = ld<T>(a.bar[1], p);
        }
        o.w0 = ld<T>(a.cur[2], p);
        o.wb0 = ld<T>(a.bar[2], p);
        if (MOIST) o.vtt = ld<T>(a.vt, (unsigned)i * NYNZ + topoff);
        if (!kLean) {
            o.mth = ld<T>(a.mprev[3], p);
            if (MOIST) {
                o.mqv = ld<T>(a.mprev[4], p);
                o.mqc = ld<T>(a.mprev[5], p);
                o.mqr = ld<T>(a.mprev[6], p);
                o.qbp = ld<T>(a.bar[4], p);
            }
        }
    };

#pragma unroll
    for (int f = 0; f < NS; f++)
#pragma unroll
        for (int d = 0; d < 6; d++) q[f][d] = ld<T>(src[f], (unsigned)(i0 - 3 + d) * NYNZ + rowoff);

    const T dtx = T(g.dtx), dty = T(g.dty), dt = T(g.dt);
    const T DTZ = T(g.dtz) * T(c.mu[kc]);
    const T rw1 = T(c.rhow[kp1]), rw0 = T(c.rhow[kc]), odr = T(c.one_d_rhou[kc]);
    const bool low = (k == 1 || k == NZ - 2);
    // The remaining column constants of this lane's level are read here, once.  A load inside the march
    // is waited for together with every load issued before it (the memory counter completes in order),
    // so one there would also wait for the next column's prefetch.  Indices are clamped for the lanes
    // outside the active levels, which never use the values.
    const T tbu = T(c.tb[kp1]) - T(c.tb[kc]), tbd = T(c.tb[kc]) - T(c.tb[km1]);
    T qbu = T(0), qbd = T(0), qbk = T(0), rk = T(1), vtden = T(0), rk64 = T(1), vtden64 = T(0);
    if (MOIST) {
        qbu = T(c.qb[kp1]) - T(c.qb[kc]);
        qbd = T(c.qb[kc]) - T(c.qb[km1]);
        qbk = T(c.qb[kc]);
        rk = T(c.rhou[kc]);
        vtden = sqrt(T(c.rhou[1]) / rk);
        if (KW == 2) {  // level 64, whose fall speed level 63 evaluates
            rk64 = T(c.rhou[KP]);
            vtden64 = sqrt(T(c.rhou[1]) / rk64);
        }
    }

    // east faces (s_east; fluxes.cpp:797-1032) of the queue's centre column with u, ubar at the column east of it
    auto east_faces = [&](T U1, T UB1, T *e) {
        T se = i5(q[0][0], q[0][1], q[0][2], q[0][3], q[0][4], q[0][5], sgn(UB1 + U1));
        T be = i5(q[1][0], q[1][1], q[1][2], q[1][3], q[1][4], q[1][5], sgn(U1));
        e[0] = U1 * be + (UB1 + U1) * se;
        if (MOIST) {
            T sv = i5(q[2][0], q[2][1], q[2][2], q[2][3], q[2][4], q[2][5], sgn(UB1 + U1));
            T bv = i5(q[3][0], q[3][1], q[3][2], q[3][3], q[3][4], q[3][5], sgn(U1));
            e[1] = U1 * bv + (UB1 + U1) * sv;
            T ub = UB1 + U1;
            e[2] = ub * i5(q[4][0], q[4][1], q[4][2], q[4][3], q[4][4], q[4][5], sgn(ub));
            e[3] = ub * i5(q[5][0], q[5][1], q[5][2], q[5][3], q[5][4], q[5][5], sgn(ub));
        }
    };
    T west[NFS], east[NFS];
    {  // west faces of column i0: the east faces of column i0-1 (compute_west_scalar, fluxes.cpp:479-501)
        const unsigned pe = (unsigned)i0 * NYNZ + rowoff;
        east_faces(ld<T>(a.cur[0], pe), ld<T>(a.bar[0], pe), west);
    }
#pragma unroll
    for (int f = 0; f < NS; f++) {
#pragma unroll
        for (int d = 0; d < 5; d++) q[f][d] = q[f][d + 1];
        q[f][5] = ld<T>(src[f], (unsigned)(i0 + 3) * NYNZ + rowoff);
    }
    Col cc{}, cs{};
    fetch_col(i0, cc);
    // two wavefronts per column: level 63 reads w, wbar of level 64 from LDS
    auto publish_w = [&]() {
        if (KW == 2 && k == KP) {
            xw[2 * r] = cc.w0;
            xw[2 * r + 1] = cc.wb0;
        }
    };
    publish_w();
    fetch_plane(i0);
    store_plane(i0, 2);
    __syncthreads();
    // Every prologue load has landed before the march starts.  Without this wait the waves that skip the
    // halo loads enter the loop with the own-column loads still counted as outstanding, and the compiler
    // then waits inside every iteration (at the first use of those registers) for most of the next
    // column's prefetch as well.
    __builtin_amdgcn_s_waitcnt(0);

    for (int i = i0; i < i1; i++) {
        const bool more = i + 1 < i1;
        const int c0 = i % 2;
        const unsigned p = (unsigned)i * NYNZ + rowoff;
        if (more) {
#pragma unroll
            for (int f = 0; f < NS; f++) qs[f] = ld<T>(src[f], (unsigned)(i + 4) * NYNZ + rowoff);
            fetch_col(i + 1, cs);
            fetch_plane(i + 1);
        }
        T mth = cc.mth, mqv = cc.mqv, mqc = cc.mqc, mqr = cc.mqr, qbp = cc.qbp;
        if (kLean) {
            mth = ld<T>(a.mprev[3], p);
            if (MOIST) {
                mqv = ld<T>(a.mprev[4], p);
                mqc = ld<T>(a.mprev[5], p);
                mqr = ld<T>(a.mprev[6], p);
                qbp = ld<T>(a.bar[4], p);
            }
        }

        east_faces(cc.u1, cc.ub1, east);
        // north faces (s_north) of row RR with v, vbar at the face
        auto north_faces = [&](int RR, T V1, T VB1, T *nf) {
#define PL(f, row) SP(c0, f)[(row) * KPW + k]
            T sn = i5(PL(0, RR - 2), PL(0, RR - 1), PL(0, RR), PL(0, RR + 1), PL(0, RR + 2), PL(0, RR + 3), sgn(VB1 + V1));
            T bn = i5(PL(1, RR - 2), PL(1, RR - 1), PL(1, RR), PL(1, RR + 1), PL(1, RR + 2), PL(1, RR + 3), sgn(V1));
            nf[0] = V1 * bn + (VB1 + V1) * sn;
            if (MOIST) {
                T sv = i5(PL(2, RR - 2), PL(2, RR - 1), PL(2, RR), PL(2, RR + 1), PL(2, RR + 2), PL(2, RR + 3), sgn(VB1 + V1));
                T bv = i5(PL(3, RR - 2), PL(3, RR - 1), PL(3, RR), PL(3, RR + 1), PL(3, RR + 2), PL(3, RR + 3), sgn(V1));
                nf[1] = V1 * bv + (VB1 + V1) * sv;
                T vb = VB1 + V1;
                nf[2] = vb * i5(PL(4, RR - 2), PL(4, RR - 1), PL(4, RR), PL(4, RR + 1), PL(4, RR + 2), PL(4, RR + 3), sgn(vb));
                nf[3] = vb * i5(PL(5, RR - 2), PL(5, RR - 1), PL(5, RR), PL(5, RR + 1), PL(5, RR + 2), PL(5, RR + 3), sgn(vb));
            }
#undef PL
        };
        T nf[NFS];
        north_faces(RO, cc.vn, cc.vbn, nf);
#pragma unroll
        for (int f = 0; f < NFS; f++) nface[(f * (TJ + 1) + r + 1) * KPW + k] = nf[f];
        if (r == 0) {
            T sf[NFS];
            north_faces(RO - 1, cc.vs, cc.vbs, sf);
#pragma unroll
            for (int f = 0; f < NFS; f++) nface[(f * (TJ + 1)) * KPW + k] = sf[f];
        }
        // top faces (s_top) of this level from the own column of the plane; w, wbar at k+1 by lane exchange
        T W1 = lane_next(cc.w0), WB1 = lane_next(cc.wb0);
        if (KW == 2 && k == KP - 1) {
            W1 = xw[2 * r];
            WB1 = xw[2 * r + 1];
        }
        auto s_top = [&](int fs, int fb) -> T {
            const T *S = SP(c0, fs) + RO * KPW, *B = SP(c0, fb) + RO * KPW;
            T st_, bt;
            if (low) {
                st_ = i2(S[kc], S[kp1]);
                bt = i2(B[kc], B[kp1]);
            } else {
                st_ = i3(S[km1], S[kc], S[kp1], S[kp2], sgn(WB1 + W1));
                bt = i3(B[km1], B[kc], B[kp1], B[kp2], sgn(W1));
            }
            return W1 * bt + (WB1 + W1) * st_;
        };
        T tf[NFS];
        tf[0] = s_top(0, 1);
        T vtk = T(0), qrt = T(0);
        if (MOIST) {
            tf[1] = s_top(2, 3);
            {
                const T *S = SP(c0, 4) + RO * KPW;
                T wb = WB1 + W1;
                T tq = low ? i2(S[kc], S[kp1]) : i3(S[km1], S[kc], S[kp1], S[kp2], sgn(wb));
                tf[2] = wb * tq;
            }
            // fall speed of this level (kessler.cpp:222-252); level 0 stores level 1's, NZ-1 keeps the array's
            auto fall = [&](T qr, T rho, T den) -> T {
                T v = T(0);
                if (qr > T(1.0e-12)) {
                    T qrr = fmax(qr * T(0.001) * rho, T(0));
                    v = T(36.34) * (pow(qrr, T(0.1364))) * den;
                }
                return v;
            };
            if (k >= 1 && k <= NZ - 2) {
                vtk = fall(q[5][2], rk, vtden);
            } else if (k == NZ - 1) {
                vtk = cc.vtt;
            }
            // rain top face with the fall speed (fluxes.cpp:734, 780; 966-1030)
            const T *QRr = SP(c0, 5) + RO * KPW;
            const T v1l = lane_read(vtk, 1);
            const T vte = k == 0 ? v1l : vtk;
            T vk1 = lane_next(vte);
            // two wavefronts per column: level 63 evaluates level 64's fall speed itself (NZ >= 66, so
            // level 64 is an interior level), from the same plane value the other wavefront uses
            if (KW == 2 && k == KP - 1) vk1 = fall(QRr[KP], rk64, vtden64);
            if (k == 0) {
                T t0 = T(0.5) * (QRr[0] + QRr[1]);
                qrt = -T(0.5) * (v1l + v1l) * t0;
            } else {
                T wb = WB1 + W1;
                T tq;
                if (kc == 1 || kc == NZ - 2) tq = i2(QRr[kc], QRr[kp1]);
                else tq = i3(QRr[km1], QRr[kc], QRr[kp1], QRr[kp2], sgn(WB1 + W1 - T(0.5) * (vte + vk1)));
                qrt = (wb - T(0.5) * (vk1 + vte)) * tq;
            }
        }
        if (k == 0) {
#pragma unroll
            for (int f = 0; f < NFS; f++) tf[f] = T(0);
        }
        T bot[NFS];
#pragma unroll
        for (int f = 0; f < NFS; f++) bot[f] = lane_prev(tf[f]);
        T qrb = MOIST ? lane_prev(qrt) : T(0);
        if (KW == 2 && k == KP - 1) {
#pragma unroll
            for (int f = 0; f < NFS; f++) xb[(NFS + 1) * r + f] = tf[f];
            xb[(NFS + 1) * r + NFS] = qrt;
        }
        __syncthreads();
        if (KW == 2 && k == KP) {
#pragma unroll
            for (int f = 0; f < NFS; f++) bot[f] = xb[(NFS + 1) * r + f];
            if (MOIST) qrb = xb[(NFS + 1) * r + NFS];
        }
        T oth = T(0), oqv = T(0), oqc = T(0), oqr = T(0);  // this level's final values (active levels)
        if (active) {
            T south[NFS];
#pragma unroll
            for (int f = 0; f < NFS; f++) south[f] = nface[(f * (TJ + 1) + r) * KPW + k];
            // Z_ADVECT_BASE (advection.cpp:16-18) with the base-state differences du = b[k+1] - b[k],
            // dd = b[k] - b[k-1] of this level
            auto zbase = [&](T du, T dd) -> T { return (T(0.50) * (rw1 * W1 * du + rw0 * cc.w0 * dd) * odr); };
            {  // theta (advection.cpp:152-168)
                T tend = -dtx * (east[0] - west[0]) - dty * (nf[0] - south[0]) - DTZ * (rw1 * tf[0] - rw0 * bot[0]) * odr -
                         DTZ * zbase(tbu, tbd);
                oth = mth + step * tend;
            }
            if (MOIST) {  // advection.cpp:197-253
                T tq = -dtx * (east[1] - west[1]) - dty * (nf[1] - south[1]) - DTZ * (rw1 * tf[1] - rw0 * bot[1]) * odr -
                       DTZ * zbase(qbu, qbd);
                T qv = mqv + step * tq;
                T tc = -dtx * (east[2] - west[2]) - dty * (nf[2] - south[2]) - DTZ * (rw1 * tf[2] - rw0 * bot[2]) * odr;
                T qc = mqc + step * tc;
                T tr = -dtx * (east[3] - west[3]) - dty * (nf[3] - south[3]) - DTZ * (rw1 * qrt - rw0 * qrb) * odr;
                T qrv = mqr + step * tr;
                // zero_moisture for the advected points (microphysics.cpp:677-698)
                if (qc < T(0)) qc = T(0);
                if (qrv < T(0)) qrv = T(0);
                if (qv + qbp + qbk < T(0)) qv = -(qbp + qbk);
                oqv = qv;
                oqc = qc;
                oqr = qrv;
            }
        }
        // the lid of upper_lower_boundaries (boundaries.cpp:304-312) for this column: level 0 takes level 1,
        // NZ-1 takes NZ-2.  These are the final values of an interior column (no sponge or periodic copy
        // touches it); the boundary epilogue redoes the columns it rewrites whole.  The sources of the
        // periodic copy are left to it: halo column 0 receives column NX-6's values from before the lid.
        // A lane writes either its own level (active levels) or a lid level (levels 0 and NZ-1).
        unsigned ps = p;
        const bool lid = j < a.jh && i >= 6 && i < g.NX - 6 && (k == 0 || k == NZ - 1);
        {
            const int lz = (NZ - 2) & (KP - 1);  // lane of level NZ-2 in its wavefront
            const int src = k == 0 ? 1 : lz;
            const T t1 = lane_read(oth, 1), tz = lane_read(oth, lz);
            T v1 = T(0), vz = T(0), c1 = T(0), cz = T(0), r1 = T(0), rz = T(0);
            if (MOIST) {
                v1 = lane_read(oqv, 1);
                vz = lane_read(oqv, lz);
                c1 = lane_read(oqc, 1);
                cz = lane_read(oqc, lz);
                r1 = lane_read(oqr, 1);
                rz = lane_read(oqr, lz);
            }
            if (lid) {
                const bool bot_lvl = src == 1;
                ps = (unsigned)i * NYNZ + (unsigned)(jc * NZ + k);
                oth = bot_lvl ? t1 : tz;
                oqv = bot_lvl ? v1 : vz;
                oqc = bot_lvl ? c1 : cz;
                oqr = bot_lvl ? r1 : rz;
            }
        }
        auto store_column = [&]() {
            if (active || lid) {
                st<T>(a.out[3], ps, oth);
                if (MOIST) {
                    st<T>(a.out[4], ps, oqv);
                    st<T>(a.out[5], ps, oqc);
                    st<T>(a.out[6], ps, oqr);
                }
            }
            if (MOIST && active) {
                st<T>(a.vt, p, vtk);
                if (k == 1) st<T>(a.vt, p - 1, vtk);
            }
        };
        if (kLean) store_column();
#pragma unroll
        for (int f = 0; f < NFS; f++) west[f] = east[f];
        if (more) {
            store_plane(i + 1, 3);
#pragma unroll
            for (int f = 0; f < NS; f++) {
#pragma unroll
                for (int d = 0; d < 5; d++) q[f][d] = q[f][d + 1];
                q[f][5] = qs[f];
            }
            cc = cs;
            publish_w();
        }
        // drained on every path (see the momentum kernel)
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        // FP32: this column's stores go out after the next column's planes are written.  The memory
        // counter completes loads and stores in order, so a store issued before that point would hold up
        // the wait for the next column's prefetched values; here it completes during the next column.
        if (!kLean) store_column();
    }
}

template <typename T, int TJ, bool MOIST, int KW>
__global__ void __launch_bounds__(KP *KW *TJ) k_adv_momentum(Geo g, Cols c, StageArgs a, T step) {
    extern __shared__ unsigned char smem_raw[];
    adv_momentum_tile<T, TJ, MOIST, KW>(g, c, a, step, xcd_tile(blockIdx.x, gridDim.x), smem_raw);
}

template <typename T, int TJ, bool MOIST, int KW>
__global__ void __launch_bounds__(KP *KW *TJ) k_adv_scalars(Geo g, Cols c, StageArgs a, T step) {
    extern __shared__ unsigned char smem_raw[];
    adv_scalars_tile<T, TJ, MOIST, KW>(g, c, a, step, xcd_tile(blockIdx.x, gridDim.x), smem_raw);
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

template <typename T, int TJ, bool MOIST, int KW>
int momentum_t(Ctx *c, const StageArgs &a0, T step, int ti) {
    StageArgs a = a0;
    a.ti = ti;
    a.nbj = (a.jh - a.jl + TJ - 1) / TJ;
    const int nbi = (a.ih - a.il + ti - 1) / ti;
    const size_t sh = MomLds<T, TJ, KW>::bytes;
    Timed _t(c, "adv_momentum", (double)(a.ih - a.il) * (a.jh - a.jl) * (c->geo.NZ - 2));
    hipLaunchKernelGGL((k_adv_momentum<T, TJ, MOIST, KW>), dim3(a.nbj * nbi), dim3(KP * KW, TJ), sh, c->stream, c->geo, c->cols, a,
                       step);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T, int TJ, bool MOIST, int KW>
int scalars_t(Ctx *c, const StageArgs &a0, T step, int ti) {
    StageArgs a = a0;
    a.ti = ti;
    a.nbj = (a.jh - a.jl + TJ - 1) / TJ;
    const int nbi = (a.ih - a.il + ti - 1) / ti;
    const size_t sh = ScalLds<T, TJ, MOIST, KW>::bytes;
    Timed _t(c, "adv_scalars", (double)(a.ih - a.il) * (a.jh - a.jl) * (c->geo.NZ - 2));
    hipLaunchKernelGGL((k_adv_scalars<T, TJ, MOIST, KW>), dim3(a.nbj * nbi), dim3(KP * KW, TJ), sh, c->stream, c->geo, c->cols, a,
                       step);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

template <typename T, bool MOIST>
int stage_t(Ctx *c, double step, int il, int ih, int jl, int jh, int parts) {
    StageArgs a{};
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
    const int mtj = fused_shape(c, DYN_TUNE_MOMENTUM_TJ), mti = fused_shape(c, DYN_TUNE_MOMENTUM_TI);
    const int stj = fused_shape(c, DYN_TUNE_SCALAR_TJ), sti = fused_shape(c, DYN_TUNE_SCALAR_TI);
    int rc = DYN_OK;
    if (fused_waves(c) == 2) {  // two wavefronts per column: 2 or 4 rows per block
        if (parts & kStageMomentum) {
            if (mtj == 2) rc = momentum_t<T, 2, MOIST, 2>(c, a, T(step), mti);
            else rc = momentum_t<T, 4, MOIST, 2>(c, a, T(step), mti);
        }
        if (rc || !(parts & kStageScalars)) return rc;
        if (stj == 2) return scalars_t<T, 2, MOIST, 2>(c, a, T(step), sti);
        return scalars_t<T, 4, MOIST, 2>(c, a, T(step), sti);
    }
    if (parts & kStageMomentum) {
        if (mtj == 2) rc = momentum_t<T, 2, MOIST, 1>(c, a, T(step), mti);
        else if (mtj == 4) rc = momentum_t<T, 4, MOIST, 1>(c, a, T(step), mti);
        else if (mtj == 16) rc = momentum_t<T, 16, MOIST, 1>(c, a, T(step), mti);
        else rc = momentum_t<T, 8, MOIST, 1>(c, a, T(step), mti);
    }
    if (rc || !(parts & kStageScalars)) return rc;
    if (stj == 2) return scalars_t<T, 2, MOIST, 1>(c, a, T(step), sti);
    if (stj == 4) return scalars_t<T, 4, MOIST, 1>(c, a, T(step), sti);
    if (stj == 12) return scalars_t<T, 12, MOIST, 1>(c, a, T(step), sti);
    return scalars_t<T, 8, MOIST, 1>(c, a, T(step), sti);
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
    // Rayleigh damping after the boundary update (end of a step): levels >= ray_level, coefficients
    // [u v theta][k] then [w][k]; ray_level >= NZ: none
    const double *ray = nullptr;
    int ray_level = 1 << 30;
};

template <typename T>
__device__ __forceinline__ T sponge_coef(const Geo &g, const PostArgs &a, int j) {
    if (j < a.js) return T(a.cs[j]);
    if (j >= g.NY - a.jn) return T(a.cn[j - (g.NY - a.jn)]);
    return T(-1);  // no sponge
}

template <typename T>
// damp: rayleigh_damping (damping.cpp:1148-1177) of the written values with coefficient cd on the
// columns it covers, i in [1, NX-1) (the caller checks j and k)
__device__ __forceinline__ void write_col(const Geo &g, T *A, int isrc, int j, int k, T fin, T prelid, bool damp = false,
                                          T cd = T(0)) {
    long long s = g.NYNZ;
    long long p = (long long)isrc * s + (long long)j * g.NZ + k;
    auto dv = [&](int i, T v) { return damp && i >= 1 && i < g.NX - 1 ? v - cd * v : v; };
    A[p] = dv(isrc, fin);
    const int shift = g.NX - 6;
    if (isrc >= shift && isrc < shift + 3) {
        int ih = isrc - shift;  // halo 0, 1, 2
        A[p - (long long)shift * s] = dv(ih, ih == 0 ? prelid : fin);
    }
    if (isrc >= 3 && isrc < 6) A[p + (long long)shift * s] = dv(isrc + shift, fin);
}

// PGF (solver.cpp:123-135) + sponge/periodic/lid for UP, VP and WP
template <typename T>
__device__ __forceinline__ void post_uvw_col(const Geo &g, const Cols &c, const PostArgs &a, T step, bool pgf, int jb, int isrc) {
    const int k = threadIdx.x;
    const int j = jb * blockDim.y + threadIdx.y;
    if (j >= g.NY) return;
    const int NZ = g.NZ;
    const int kc = k < NZ ? k : NZ - 1;
    const long long p = (long long)isrc * g.NYNZ + (long long)j * NZ + kc;
    T *UP = (T *)a.f[0], *VP = (T *)a.f[1], *WP = (T *)a.f[2];
    const T *PI = (const T *)a.pi;
    T u = UP[p], v = VP[p], w = WP[p];
    // every load of the column goes out before the arithmetic: the neighbours p - NYNZ, p - NZ, p - 1
    // lie inside the array for every source column (isrc >= 3), used or not
    const T pc = PI[p], pw = PI[p - g.NYNZ], psn = PI[p - NZ], pb = PI[p - 1];
    const double mwk = c.mw[kc];
    if (pgf && j >= 1 && j <= g.NY - 2 && kc >= 1 && kc <= NZ - 2) {
        u -= step * (T(g.dtx) * (pc - pw));
        v -= step * (T(g.dty) * (pc - psn));
        if (kc >= 2) w -= step * ((T(g.dtz) * T(mwk)) * (pc - pb));
    }
    T coef = sponge_coef<T>(g, a, j);
    if (coef >= T(0) && kc >= 1) {
        u -= coef * u;
        v -= coef * v;
        w -= coef * w;
    }
    // lid from lanes 1 and NZ-2 (j >= 1; every source column has i >= 1)
    const int lz = (NZ - 2) & (KP - 1);  // lane of level NZ-2 in its wavefront
    T u1 = lane_read(u, 1), uz = lane_read(u, lz), v1 = lane_read(v, 1), vz = lane_read(v, lz);
    if (k >= NZ) return;
    T uf = u, vf = v, wf = w;
    if (j >= 1) {
        if (k == 0) { uf = u1; vf = v1; }
        if (k == NZ - 1) { uf = uz; vf = vz; }
        if (k == 0 || k == 1 || k == NZ - 1) wf = T(0);
    }
    // Rayleigh damping after the boundary update (j in [1, NY-1), k >= level)
    const bool dm = k >= a.ray_level && j >= 1 && j < g.NY - 1;
    const T cu = dm ? T(a.ray[k]) : T(0), cw = dm ? T(a.ray[NZ + k]) : T(0);
    write_col(g, UP, isrc, j, k, uf, u, dm, cu);
    write_col(g, VP, isrc, j, k, vf, v, dm, cu);
    write_col(g, WP, isrc, j, k, wf, w, dm, cw);
}

// sponge/periodic/lid for THP (and QVP, QCP, QRP with zero_moisture outside the advected box); only the
// columns a sponge row, a ghost row or the periodic copy touches are rewritten whole, the others get the
// two lid levels
// lid_interior: also set the lid of the interior columns (the stage advection already did after a stage;
// after Kessler it must be redone)
template <typename T, bool MOIST>
__device__ __forceinline__ void post_scalars_col(const Geo &g, const Cols &c, const PostArgs &a, bool clip, bool lid_interior, int j,
                                                 int isrc) {
    const int k = threadIdx.x;
    if (j >= g.NY) return;
    const int NZ = g.NZ;
    const int kc = k < NZ ? k : NZ - 1;
    const long long col = (long long)isrc * g.NYNZ + (long long)j * NZ;
    const T coef = sponge_coef<T>(g, a, j);
    const bool full = coef >= T(0) || j < 3 || j >= g.NY - 3 || isrc < 6 || isrc >= g.NX - 6;
    constexpr int NFLD = MOIST ? 4 : 1;
    if (!full) {  // interior column: lid only (values at 1 and NZ-2 are final already)
        // with Rayleigh damping (end of a step) theta is rewritten from level ray_level up as well: the
        // lid from the undamped levels 1 and NZ-2, then the damping (j in [3, NY-3) here)
        const bool ray = a.ray_level < NZ;
        if (lid_interior && j >= 1 && (k == 0 || k == NZ - 1)) {
            int ks = k == 0 ? 1 : NZ - 2;
#pragma unroll
            for (int f = ray ? 1 : 0; f < NFLD; f++) {
                T *A = (T *)a.f[3 + f];
                A[col + k] = A[col + ks];
            }
        }
        if (ray) {
            T *A = (T *)a.f[3];
            const T x = A[col + kc];
            const T x1 = lane_read(x, 1), xz = lane_read(x, (NZ - 2) & (KP - 1));
            const bool lid = lid_interior && (k == 0 || k == NZ - 1);
            if (k < NZ && (lid || k >= a.ray_level)) {
                T fin = k == 0 && lid ? x1 : (k == NZ - 1 && lid ? xz : x);
                if (k >= a.ray_level) fin = fin - T(a.ray[k]) * fin;
                A[col + k] = fin;
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
        T x1 = lane_read(x, 1), xz = lane_read(x, (NZ - 2) & (KP - 1));
        if (k < NZ) {
            T fin = x;
            if (j >= 1) {
                if (k == 0) fin = x1;
                if (k == NZ - 1) fin = xz;
            }
            // Rayleigh damping of theta after its boundary update (j in [1, NY-1), k >= level)
            const bool dm = f == 0 && k >= a.ray_level && j >= 1 && j < g.NY - 1;
            write_col(g, A, isrc, j, k, fin, x, dm, dm ? T(a.ray[k]) : T(0));
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_post_uvw(Geo g, Cols c, PostArgs a, T step, bool pgf) {
    post_uvw_col<T>(g, c, a, step, pgf, blockIdx.x, blockIdx.y + 3);
}
template <typename T, bool MOIST>
__global__ void __launch_bounds__(256) k_post_scalars(Geo g, Cols c, PostArgs a, bool clip, bool lid_interior) {
    post_scalars_col<T, MOIST>(g, c, a, clip, lid_interior, blockIdx.x * blockDim.y + threadIdx.y, blockIdx.y + 3);
}

// Source columns post_scalars_col rewrites whole when the interior lids are already set (after a
// stage): rows j < rl and j >= NY - rh (sponge and ghost rows) for every isrc in [3, NX-3), then, on the
// other rows, the six periodic source columns.  One wavefront per column, item = its index in that list.
struct ShellCols {
    int rl, rh, nb, ni;  // low rows, high rows, rows between, source columns (NX - 6)
    int ncb;             // periodic source columns (6, or every column on a narrow grid)
    int total;
};
inline ShellCols shell_cols(const Ctx *c) {
    ShellCols s;
    const int NY = c->g.ny, ni = c->g.nx - 6;
    s.rl = std::min(NY, std::max(c->g.sponge_south, 3));
    s.rh = std::min(NY - s.rl, std::max(c->g.sponge_north, 3));
    s.nb = NY - s.rl - s.rh;
    s.ni = ni;
    s.ncb = ni <= 6 ? ni : 6;
    s.total = (s.rl + s.rh) * ni + s.ncb * s.nb;
    return s;
}
template <typename T, bool MOIST>
__global__ void __launch_bounds__(256) k_post_scalars_shell(Geo g, Cols c, PostArgs a, bool clip, ShellCols s) {
    const int item = blockIdx.x * blockDim.y + threadIdx.y;
    if (item >= s.total) return;
    const int na = (s.rl + s.rh) * s.ni;
    int j, isrc;
    if (item < na) {
        const int r = item / s.ni;
        isrc = 3 + item % s.ni;
        j = r < s.rl ? r : g.NY - s.rh + (r - s.rl);
    } else {
        const int q = item - na, cb = q / s.nb;
        j = s.rl + q % s.nb;
        isrc = s.ni <= 6 ? 3 + cb : (cb < 3 ? 3 + cb : g.NX - 6 + (cb - 3));
    }
    post_scalars_col<T, MOIST>(g, c, a, clip, false, j, isrc);
}

// ---- batched over the members of an ensemble: one block per work item of the concatenated work list ----
__device__ __forceinline__ StageArgs ens_stage_args(const EnsMember &e, const EnsRoles &ro, int ti, int nbj) {
    StageArgs a;
    for (int f = 0; f < 7; f++) {
        a.cur[f] = e.buf[f][ro.r[f][DYN_ROLE_CUR]];
        a.mprev[f] = e.buf[f][ro.r[f][DYN_ROLE_M]];
        a.out[f] = e.buf[f][ro.r[f][DYN_ROLE_P]];
    }
    for (int f = 0; f < 5; f++) a.bar[f] = e.bar[f];
    a.vt = e.vt;
    a.il = 3;
    a.ih = e.g.NX - 3;
    a.jl = 3;
    a.jh = e.g.NY - 3;
    a.ti = ti;
    a.nbj = nbj;
    return a;
}
__device__ __forceinline__ PostArgs ens_post_args(const EnsMember &e, const EnsRoles &ro) {
    PostArgs a;
    for (int f = 0; f < 7; f++) a.f[f] = e.buf[f][ro.r[f][DYN_ROLE_P]];
    a.pi = e.pi;
    a.qb = e.bar[4];
    a.cs = e.sponge_s;
    a.cn = e.sponge_n;
    a.js = e.js;
    a.jn = e.jn;
    a.zl = 1;
    a.zh = e.g.NX - 1;
    a.zjl = 1;
    a.zjh = e.g.NY - 1;
    return a;
}

template <typename T, int TJ, bool MOIST>
__global__ void __launch_bounds__(KP *TJ) k_ens_momentum(const EnsMember *mem, int nmem, EnsRoles ro, T step, int ti, int total) {
    extern __shared__ unsigned char smem_raw[];
    const int item = blockIdx.x;
    const EnsMember &e = mem[ens_member(mem, nmem, item, &EnsMember::mom0)];
    const StageArgs a = ens_stage_args(e, ro, ti, e.nbj_mom);
    adv_momentum_tile<T, TJ, MOIST>(e.g, e.c, a, step, item - e.mom0, smem_raw);
}

template <typename T, int TJ, bool MOIST>
__global__ void __launch_bounds__(KP *TJ) k_ens_scalars(const EnsMember *mem, int nmem, EnsRoles ro, T step, int ti, int total) {
    extern __shared__ unsigned char smem_raw[];
    const int item = blockIdx.x;
    const EnsMember &e = mem[ens_member(mem, nmem, item, &EnsMember::scal0)];
    const StageArgs a = ens_stage_args(e, ro, ti, e.nbj_scal);
    adv_scalars_tile<T, TJ, MOIST>(e.g, e.c, a, step, item - e.scal0, smem_raw);
}

template <typename T>
__global__ void __launch_bounds__(256) k_ens_post(const EnsMember *mem, int nmem, EnsRoles ro, T step, bool pgf, bool scalars, int total) {
    const int item = blockIdx.x;
    const EnsMember &e = mem[ens_member(mem, nmem, item, &EnsMember::post0)];
    const PostArgs a = ens_post_args(e, ro);
    const int local = item - e.post0, nb = (e.g.NY + 3) / 4;
    post_uvw_col<T>(e.g, e.c, a, step, pgf, local % nb, local / nb + 3);
    if (scalars) post_scalars_col<T, false>(e.g, e.c, a, false, !pgf, (local % nb) * blockDim.y + threadIdx.y, local / nb + 3);
}

// advance_outer's memcpy(M <- P) for u, v, w, theta of every member; work items are 1024-element chunks
template <typename T>
__global__ void __launch_bounds__(256) k_ens_copy(const EnsMember *mem, int nmem, EnsRoles ro, int total) {
    {
        const int item = blockIdx.x;
        const EnsMember &e = mem[ens_member(mem, nmem, item, &EnsMember::cp0)];
        const long long n = (long long)e.g.NX * e.g.NYNZ;
        const long long per = (n + 1023) / 1024;
        const int local = item - e.cp0;
        const int fi = local / (int)per;
        const int f = fi == 0 ? DYN_U : fi == 1 ? DYN_V : fi == 2 ? DYN_W : DYN_TH;
        T *dst = (T *)e.buf[f][ro.r[f][DYN_ROLE_M]];
        const T *src = (const T *)e.buf[f][ro.r[f][DYN_ROLE_P]];
        const long long base = (long long)(local % (int)per) * 1024;
        for (long long p = base + threadIdx.x; p < base + 1024 && p < n; p += blockDim.x) dst[p] = src[p];
    }
}

template <typename T>
int post_t(Ctx *c, double step, bool pgf, bool scalars, bool clip, bool uvw, int ray_level, bool interior_done) {
    PostArgs a{};
    if (ray_level >= 0 && ray_level < c->g.nz) {
        a.ray = c->raycoef;
        a.ray_level = ray_level;
    }
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
    // one wavefront per column, two when NZ > 64; 256 threads per block
    const int kw = c->g.nz > KP ? 2 : 1, rows = 4 / kw;
    dim3 blk(KP * kw, rows), grd((c->g.ny + rows - 1) / rows, c->g.nx - 6);
    if (uvw) {
        Timed _t(c, "post_uvw", (double)c->npts);
        hipLaunchKernelGGL(k_post_uvw<T>, grd, blk, 0, c->stream, c->geo, c->cols, a, T(step), pgf);
        DYN_LAUNCH_CHECK();
    }
    if (scalars) {
        Timed _t(c, "post_scalars", (double)c->npts);
        // after a stage the advection kernel has set the interior lids; at the end of the step (no
        // gradient force) Kessler has changed the interior, so they are redone
        const bool lid_interior = !pgf;
        if (!lid_interior || interior_done) {  // only the sponge rows, ghost rows and periodic source columns change
            const ShellCols sc = shell_cols(c);
            const dim3 gs((unsigned)std::max(1, (sc.total + rows - 1) / rows));
            if (c->g.microphysics)
                hipLaunchKernelGGL((k_post_scalars_shell<T, true>), gs, blk, 0, c->stream, c->geo, c->cols, a, clip, sc);
            else hipLaunchKernelGGL((k_post_scalars_shell<T, false>), gs, blk, 0, c->stream, c->geo, c->cols, a, clip, sc);
        } else if (c->g.microphysics) {
            hipLaunchKernelGGL((k_post_scalars<T, true>), grd, blk, 0, c->stream, c->geo, c->cols, a, clip, lid_interior);
        } else {
            hipLaunchKernelGGL((k_post_scalars<T, false>), grd, blk, 0, c->stream, c->geo, c->cols, a, clip, lid_interior);
        }
        DYN_LAUNCH_CHECK();
    }
    return DYN_OK;
}

}  // namespace

// Wavefronts per column of the stage kernels: one up to NZ = 64; two for NZ in [66, 128], where the lid's
// source levels 1 and NZ-2 sit in the wavefronts of the levels they feed (0 and NZ-1); 0 otherwise
int fused_waves(const Ctx *c) {
    const int nz = c->g.nz;
    return nz <= KP ? 1 : (nz >= KP + 2 && nz <= 2 * KP ? 2 : 0);
}

// ... and 32-bit byte offsets into every field
// the fused stage kernels carry no turbulence: with it on, steps take the per-operator sequence
bool fused_supported(const Ctx *c) {
    return !c->turb.on && fused_waves(c) > 0 && (double)c->npts * (double)c->elt < 4294967296.0;
}

namespace {
template <typename T, int KW>
size_t stage_lds_t(bool momentum, bool moist, int tj) {
    if (tj == 2) return momentum ? MomLds<T, 2, KW>::bytes : moist ? ScalLds<T, 2, true, KW>::bytes : ScalLds<T, 2, false, KW>::bytes;
    if (tj == 4) return momentum ? MomLds<T, 4, KW>::bytes : moist ? ScalLds<T, 4, true, KW>::bytes : ScalLds<T, 4, false, KW>::bytes;
    if (tj == 12) return moist ? ScalLds<T, 12, true, KW>::bytes : ScalLds<T, 12, false, KW>::bytes;  // scalar kernel only
    if (tj == 16) return MomLds<T, 16, KW>::bytes;  // momentum kernel only
    return momentum ? MomLds<T, 8, KW>::bytes : moist ? ScalLds<T, 8, true, KW>::bytes : ScalLds<T, 8, false, KW>::bytes;
}
// LDS bytes per block of a stage kernel with tj rows
size_t stage_lds(const Ctx *c, bool momentum, int tj) {
    const bool moist = c->g.microphysics != 0, two = fused_waves(c) == 2;
    if (c->prec == DYN_FP32) return two ? stage_lds_t<float, 2>(momentum, moist, tj) : stage_lds_t<float, 1>(momentum, moist, tj);
    return two ? stage_lds_t<double, 2>(momentum, moist, tj) : stage_lds_t<double, 1>(momentum, moist, tj);
}
}  // namespace

// launch-shape knobs of the stage kernels: rows per block (instantiated widths: 2, 4, 8 with one wavefront
// per column, and 16 for the momentum kernel, 12 for the scalar kernel; 2, 4 with two; within the LDS)
// and march length
bool fused_shape_ok(const Ctx *c, int param, int value) {
    if (param == DYN_TUNE_MOMENTUM_TJ || param == DYN_TUNE_SCALAR_TJ) {
        const bool one = fused_waves(c) != 2, mom = param == DYN_TUNE_MOMENTUM_TJ;
        const bool inst = value == 2 || value == 4 || (one && (value == 8 || value == (mom ? 16 : 12)));
        return inst && stage_lds(c, param == DYN_TUNE_MOMENTUM_TJ, value) <= c->pp.lds_max;
    }
    if (param == DYN_TUNE_MOMENTUM_TI || param == DYN_TUNE_SCALAR_TI) return value >= 2 && value <= 512;
    if (param == DYN_TUNE_PRESSURE_JPASS) return value == 2 || (value == 1 && c->pp.fused_y) || (value == 3 && c->pp.half_y);
    if (param == DYN_TUNE_STAGE_STREAMS) return value >= 1 && value <= 4;
    return false;
}

namespace {
// The FP32 scalar kernel on one stream (no pressure pass beside it): 12 rows per block (three waves per
// SIMD at its register count, one block per CU) and march lengths that make one round of blocks fill the
// device.  Measured at 256x256x64 moist: 92.7 us against 104.2 us for 8 rows x 16 columns, one-stream
// step 0.956 against 0.990 ms.  Beside the pressure solve 8 rows stay faster (the scalar blocks must
// leave room for the solve's).
bool scalar_one_round(const Ctx *c) {
    if (c->prec != DYN_FP32 || fused_waves(c) != 1 || c->pp.cus <= 0) return false;
    if (c->tune[DYN_TUNE_STAGE_STREAMS] != 1) return false;  // one stream only when asked for
    return fused_shape_ok(c, DYN_TUNE_SCALAR_TJ, 12);
}
}  // namespace

int fused_shape(const Ctx *c, int param) {
    if (c->tune[param]) return c->tune[param];
    const bool f32 = c->prec == DYN_FP32;
    // FP64: 8 rows per block when that still gives two blocks per CU at 16 columns per block (measured
    // 8 % faster at 256x256x64), else 4
    const int rows8 = ((c->g.ny - 6 + 7) / 8) * ((c->g.nx - 6 + 15) / 16);
    const int tj = f32 || rows8 >= 512 ? 8 : 4;
    switch (param) {
    case DYN_TUNE_MOMENTUM_TJ:
    case DYN_TUNE_SCALAR_TJ:
        // two wavefronts per column: blocks of 512 threads (4 rows) where the LDS allows and the moist
        // FP64 scalar kernel keeps its registers (at 4 rows it spills), else 2 rows
        if (fused_waves(c) == 2) {
            if (!f32 && param == DYN_TUNE_SCALAR_TJ && c->g.microphysics) return 2;
            return fused_shape_ok(c, param, 4) ? 4 : 2;
        }
        if (param == DYN_TUNE_SCALAR_TJ && scalar_one_round(c)) return 12;
        return tj;
    case DYN_TUNE_MOMENTUM_TI: return fused_waves(c) == 2 ? 32 : 16;  // 512x512x80: 32 measured 3 % faster
    // The scalar kernel beside the pressure solve (fork after the momentum kernel): 24 columns, about 2 %
    // faster per step than 16 (FP32 at 256x256x64 moist and 192x192x40 dry, FP64 at 192x192x40 dry).
    // Its 352 blocks at 256x256 leave most CUs one block, beside which a j-pass block still fits; alone
    // the kernel is slower that way (165 against 120 us), so the one-stream step keeps 16, and so does
    // the later fork, where 16 measured best.
    case DYN_TUNE_SCALAR_TI:
        if (fused_waves(c) == 2) return 32;
        if (scalar_one_round(c)) {  // as many i-chunks per 12-row band as leave every block one CU, at
                                    // least 16 columns each (shorter marches spend more on the queue fill)
            const int nbj = (c->g.ny - 6 + 11) / 12, per = std::max(1, c->pp.cus / nbj);
            return std::min(512, std::max(16, (c->g.nx + per - 1) / per));
        }
        {
            const int streams = fused_shape(c, DYN_TUNE_STAGE_STREAMS);
            return streams == 2 || streams == 4 ? 24 : 16;
        }
    case DYN_TUNE_PRESSURE_JPASS: return c->pp.fused_y ? 1 : (c->pp.half_y ? 3 : 2);
    // the scalar half forked after the pressure solve's first pass where the j pass runs on half slabs
    // (FP64 256x256x64: 4 % faster); after the momentum kernel elsewhere (FP32 there: 7 % faster), in
    // FP64 with its launches queued after the first pass (192x192x40: 5 % faster)
    case DYN_TUNE_STAGE_STREAMS: return fused_shape(c, DYN_TUNE_PRESSURE_JPASS) == 3 ? 3 : (f32 ? 2 : 4);
    default: return DYN_E_ARG;
    }
}

// shapes whose LDS exceeds the device's are never launched (fused_shape_ok), so their limit stays unset
template <typename T, int TJ, int KW>
int set_lds_limits(const Ctx *c) {
    const size_t mom = MomLds<T, TJ, KW>::bytes, sm = ScalLds<T, TJ, true, KW>::bytes, sd = ScalLds<T, TJ, false, KW>::bytes;
    if (mom <= c->pp.lds_max) {
        DYN_HIP(hipFuncSetAttribute((const void *)k_adv_momentum<T, TJ, true, KW>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mom));
        DYN_HIP(hipFuncSetAttribute((const void *)k_adv_momentum<T, TJ, false, KW>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mom));
    }
    if (sm <= c->pp.lds_max)
        DYN_HIP(hipFuncSetAttribute((const void *)k_adv_scalars<T, TJ, true, KW>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    if (sd <= c->pp.lds_max)
        DYN_HIP(hipFuncSetAttribute((const void *)k_adv_scalars<T, TJ, false, KW>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sd));
    return DYN_OK;
}

template <typename T>
int set_lds_limits_all(const Ctx *c) {
    if (fused_waves(c) == 2) {
        int rc = set_lds_limits<T, 2, 2>(c);
        if (!rc) rc = set_lds_limits<T, 4, 2>(c);
        return rc;
    }
    int rc = set_lds_limits<T, 2, 1>(c);
    if (!rc) rc = set_lds_limits<T, 4, 1>(c);
    if (!rc) rc = set_lds_limits<T, 8, 1>(c);
    if (rc) return rc;
    const size_t mom = MomLds<T, 16, 1>::bytes, sm = ScalLds<T, 12, true, 1>::bytes, sd = ScalLds<T, 12, false, 1>::bytes;
    if (mom <= c->pp.lds_max) {
        DYN_HIP(hipFuncSetAttribute((const void *)k_adv_momentum<T, 16, true, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mom));
        DYN_HIP(hipFuncSetAttribute((const void *)k_adv_momentum<T, 16, false, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mom));
    }
    if (sm <= c->pp.lds_max)
        DYN_HIP(hipFuncSetAttribute((const void *)k_adv_scalars<T, 12, true, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    if (sd <= c->pp.lds_max)
        DYN_HIP(hipFuncSetAttribute((const void *)k_adv_scalars<T, 12, false, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sd));
    return DYN_OK;
}

// the stage kernels may use more than 64 KiB of LDS: raise the per-launch limit once, outside any capture
int fused_prepare(Ctx *c) { return c->prec == DYN_FP32 ? set_lds_limits_all<float>(c) : set_lds_limits_all<double>(c); }

// advect_uvw_velocity + advect_theta (+ advect_microphysics_cell and the interior zero_moisture)
// over the box [il,ih) x [jl,jh)
int launch_fused_advect(Ctx *c, double step, int il, int ih, int jl, int jh, int parts) {
    // the i-queues reach columns il-3 .. ih+2 and the planes rows jl-3 .. jh+2
    if (il < 3 || ih > c->g.nx - 3 || jl < 3 || jh > c->g.ny - 3 || il >= ih || jl >= jh) {
        set_error("fused stage: box outside [3, NX-3) x [3, NY-3)");
        return DYN_E_ARG;
    }
    const bool moist = c->g.microphysics != 0;
    if (c->prec == DYN_FP32)
        return moist ? stage_t<float, true>(c, step, il, ih, jl, jh, parts) : stage_t<float, false>(c, step, il, ih, jl, jh, parts);
    return moist ? stage_t<double, true>(c, step, il, ih, jl, jh, parts) : stage_t<double, false>(c, step, il, ih, jl, jh, parts);
}

// PGF (optional) and the boundary update of every prognostic after launch_fused_advect; stands for
// dyn_pgf + zero_moisture's remainder + dyn_apply_bc + dyn_apply_bc_micro of one stage
int launch_post_stage(Ctx *c, double step, bool pgf, bool scalars, bool clip, bool uvw, int ray_level, bool interior_done) {
    if (c->prec == DYN_FP32) return post_t<float>(c, step, pgf, scalars, clip, uvw, ray_level, interior_done);
    return post_t<double>(c, step, pgf, scalars, clip, uvw, ray_level, interior_done);
}

// the rest of zero_moisture(il,ih,jl,jh) after launch_fused_advect on the box [3,NX-3) x [3,NY-3)
int launch_zero_moisture_shell(Ctx *c, int il, int ih, int jl, int jh) {
    if (c->prec == DYN_FP32) return shell_t<float>(c, il, ih, jl, jh);
    return shell_t<double>(c, il, ih, jl, jh);
}


// ---- ensemble launchers ----------------------------------------------------------------------------
namespace {
template <typename T, int TJ>
int ens_momentum_t(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, T step, int ti, int total) {
    const size_t sh = MomLds<T, TJ>::bytes;
    const int grid = total;
    Timed _t(c0, "ens_momentum", 0.0);
    hipLaunchKernelGGL((k_ens_momentum<T, TJ, false>), dim3(grid), dim3(KP, TJ), sh, c0->stream, d, nmem, ro, step, ti, total);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}
template <typename T, int TJ>
int ens_scalars_t(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, T step, int ti, int total) {
    const size_t sh = ScalLds<T, TJ, false>::bytes;
    const int grid = total;
    Timed _t(c0, "ens_scalars", 0.0);
    hipLaunchKernelGGL((k_ens_scalars<T, TJ, false>), dim3(grid), dim3(KP, TJ), sh, c0->stream, d, nmem, ro, step, ti, total);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}
template <typename T>
int ens_stage_t(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, double step, const int *shape, int tot_mom, int tot_scal) {
    int rc;
    if (shape[0] == 2) rc = ens_momentum_t<T, 2>(c0, d, nmem, ro, T(step), shape[1], tot_mom);
    else if (shape[0] == 4) rc = ens_momentum_t<T, 4>(c0, d, nmem, ro, T(step), shape[1], tot_mom);
    else rc = ens_momentum_t<T, 8>(c0, d, nmem, ro, T(step), shape[1], tot_mom);
    if (rc) return rc;
    if (shape[2] == 2) return ens_scalars_t<T, 2>(c0, d, nmem, ro, T(step), shape[3], tot_scal);
    if (shape[2] == 4) return ens_scalars_t<T, 4>(c0, d, nmem, ro, T(step), shape[3], tot_scal);
    return ens_scalars_t<T, 8>(c0, d, nmem, ro, T(step), shape[3], tot_scal);
}
template <typename T>
int ens_post_t(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, double step, bool pgf, int total) {
    const int grid = total;
    Timed _t(c0, "ens_post", 0.0);
    hipLaunchKernelGGL(k_ens_post<T>, dim3(grid), dim3(KP, 4), 0, c0->stream, d, nmem, ro, T(step), pgf, true, total);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}
template <typename T>
int ens_copy_t(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, int total) {
    const int grid = total;
    Timed _t(c0, "ens_advance_copy", 0.0);
    hipLaunchKernelGGL(k_ens_copy<T>, dim3(grid), dim3(256), 0, c0->stream, d, nmem, ro, total);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}
template <typename T, int TJ>
int ens_lds_limits() {
    DYN_HIP(hipFuncSetAttribute((const void *)k_ens_momentum<T, TJ, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)MomLds<T, TJ>::bytes));
    DYN_HIP(hipFuncSetAttribute((const void *)k_ens_scalars<T, TJ, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)ScalLds<T, TJ, false>::bytes));
    return DYN_OK;
}
}  // namespace

int ens_prepare(bool fp32) {
    int rc = fp32 ? ens_lds_limits<float, 2>() : ens_lds_limits<double, 2>();
    if (!rc) rc = fp32 ? ens_lds_limits<float, 4>() : ens_lds_limits<double, 4>();
    if (!rc) rc = fp32 ? ens_lds_limits<float, 8>() : ens_lds_limits<double, 8>();
    return rc;
}
// the stage advection of every member (one momentum and one scalar launch); shape = {mom TJ, mom TI, scal TJ, scal TI}
int launch_ens_stage(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, double step, const int *shape, int tot_mom, int tot_scal) {
    return c0->prec == DYN_FP32 ? ens_stage_t<float>(c0, d, nmem, ro, step, shape, tot_mom, tot_scal)
                                : ens_stage_t<double>(c0, d, nmem, ro, step, shape, tot_mom, tot_scal);
}
int launch_ens_post(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, double step, bool pgf, int total) {
    return c0->prec == DYN_FP32 ? ens_post_t<float>(c0, d, nmem, ro, step, pgf, total) : ens_post_t<double>(c0, d, nmem, ro, step, pgf, total);
}
int launch_ens_copy(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, int total) {
    return c0->prec == DYN_FP32 ? ens_copy_t<float>(c0, d, nmem, ro, total) : ens_copy_t<double>(c0, d, nmem, ro, total);
}

}  // namespace dyn
