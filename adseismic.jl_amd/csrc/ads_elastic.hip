// ads_elastic.hip -- elastic staggered-grid velocity-stress stepper (4th order, CPML) for
// gfx950: forward, hand-written adjoint for d/d(rho, lambda, mu), checkpointed shot batches.
//
// One step (src/Core.jl:96-228), 0-based cells, N-S = x = i, E-W = y = j:
//   fw1 (i in [1,NX-1], j in [2,NY]): X = Dx+ vx + m1', Y = Dy- vy + m2',
//        sxx += dt ((lb+2mb) X + lb Y), syy += dt ((lb+2mb) Y + lb X),
//        lb = (lam_i + lam_{i+1})/2, mb likewise; m1' = bxh m1 + axh Dx+ vx, m2' = byf m2 + ayf Dy- vy
//   fw2 (i in [2,NX], j in [1,NY-1]): sxy += dt mb2 (Dx- vy + m3' + Dy+ vx + m4'), mb2 = (mu_j+mu_{j+1})/2
//   fw3 (i in [2,NX], j in [2,NY]):   vx += dt (Dx- sxx + m5' + Dy- sxy + m6') / rho
//   fw4 (i in [1,NX-1], j in [1,NY-1]): vy += dt (Dx+ sxy + m7' + Dy+ syy + m8') / rb,
//        rb = (rho_ij + rho_{i+1,j} + rho_{i+1,j+1} + rho_{i,j+1})/4
//   then sources: field_type[src] += srcv[n-1, s] (AddSource.cpp:57-85).
// D+ f_i = (27(f_{i+1}-f_i) - f_{i+2} + f_{i-1})/(24h), D- f_i = (27(f_i-f_{i-1}) - f_{i+1} + f_{i-2})/(24h).
// Memory variables outside their update range are 0 (makevector, Core.jl:697-699); the PML
// profile index is i-1 / j-1 (Core.jl:116-117, 146-147, 174-175, 205-206).  kappa is not
// applied (adbroadcast idx 3/4 discards its result, Core.jl:686-693).
//
// State layout: a level holds 13 arrays per shot (vx, vy, sxx, syy, sxy, m1..m8); the stress
// kernel reads level n-1 and writes sxx..sxy, m1..m4 of level n; the velocity kernel reads the
// new stresses and writes vx, vy, m5..m8 of level n.  The adjoint keeps the 13 cotangents and
// runs, per level in reverse: residual injection, velocity-phase VJP (2 passes: per-cell
// cotangents of the derivatives, then their transposed-stencil gather) and stress-phase VJP
// (2 passes).  Sigma-type sources are removed from level n before its velocity VJP (fw3/fw4
// read the stresses before injection).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "ads_common.h"
#include "ads_schedule.h"
#include "ads_stencil.h"

namespace ads {

constexpr int NF = 13;  // fields per level
enum { VX = 0, VY, SXX, SYY, SXY, M1, M2, M3, M4, M5, M6, M7, M8 };

// a level of the 13 arrays: field pointers and per-field shot strides.  The memory variables
// M1..M8 are either full grids like the fields (mb.H == 0) or kept only on the cells outside
// the box mb where every CPML profile vanishes (there they stay 0): mv() gives their index.
template <typename T> struct El {
    T *p[NF];
    int64_t st[NF];
    FrameBox mb;
    __device__ __host__ T *f(int q, int s) const { return p[q] + (int64_t)s * st[q]; }
    __device__ __host__ int64_t mv(int i, int j, int64_t ij) const { return mb.H ? mb.index(i, j) : ij; }
};
template <typename T> struct ElC {
    const T *p[NF];
    int64_t st[NF];
    FrameBox mb;
    __device__ __host__ const T *f(int q, int s) const { return p[q] + (int64_t)s * st[q]; }
    __device__ __host__ int64_t mv(int i, int j, int64_t ij) const { return mb.H ? mb.index(i, j) : ij; }
};
template <typename T> ElC<T> cst(const El<T> &a)
{
    ElC<T> c;
    for (int q = 0; q < NF; q++) c.p[q] = a.p[q], c.st[q] = a.st[q];
    c.mb = a.mb;
    return c;
}

template <typename T> struct ElCoef {
    T dt, ihx, ihy;  // 1/(24 hx), 1/(24 hy)
    const T *axf, *axh, *bxf, *bxh, *ayf, *ayh, *byf, *byh;
    // the same coefficients packed per padded row r (prow[4 (r + 2) + 0..3] = axf, axh, bxf,
    // bxh for r in [-2, NX+3], 0 off the profile range) and column (pcol[4 j + 0..3] = ayf,
    // ayh, byf, byh for j in [0, NY+1])
    const T *prow, *pcol;
    int NX, NY, P;
    // sub = 1: one MPIElastic subdomain (MPIElastic.jl:483-620): a block of (NX-2) x (NY-2)
    // cells at [2, NX-1] x [2, NY-1] inside a 2-cell zero halo; every fw updates the whole
    // block, parameters are not averaged and the profile index is i-2 / j-2.
    int sub;
};

// update range of fw1..fw4 (Core.jl:100-107, 132-139, 162-169, 190-197), or the block
template <typename T> __device__ inline bool fw_in(const ElCoef<T> &c, int fw, int i, int j)
{
    if (c.sub) return i >= 2 && i <= c.NX - 1 && j >= 2 && j <= c.NY - 1;
    switch (fw) {
    case 1: return i >= 1 && i <= c.NX - 1 && j >= 2 && j <= c.NY;
    case 2: return i >= 2 && i <= c.NX && j >= 1 && j <= c.NY - 1;
    case 3: return i >= 2 && i <= c.NX && j >= 2 && j <= c.NY;
    default: return i >= 1 && i <= c.NX - 1 && j >= 1 && j <= c.NY - 1;
    }
}

// ------------------------------- forward -------------------------------------------------
// The padded grid is split once per plan into an interior rectangle, where every update range
// holds and the CPML coefficients are the identity (a = 0, b = 1, so the memory variables stay
// 0 and are neither read nor written), and the frame around it, which runs the full per-cell
// scheme below.  The memory variables of rectangle cells are never read, so their slot
// values are left as they are.

// one cell of fw1/fw2 (level n-1 -> n), memory variables included
template <typename T>
__device__ inline void el_stress_cell(const ElC<T> &I, const El<T> &O, const T *__restrict__ lam,
                                      const T *__restrict__ mu, const ElCoef<T> &c, int s, int i,
                                      int j)
{
    const int P = c.P;
    const int64_t ij = (int64_t)i * P + j;
    const T *__restrict__ vx = I.f(VX, s);
    const T *__restrict__ vy = I.f(VY, s);
    T sxx = I.f(SXX, s)[ij], syy = I.f(SYY, s)[ij], sxy = I.f(SXY, s)[ij];
    T m1 = 0, m2 = 0, m3 = 0, m4 = 0;
    const int64_t mi = I.mv(i, j, ij), mo = O.mv(i, j, ij);  // memory variables (< 0: all 0)
    auto mold = [&](int q) { return mi >= 0 ? I.f(q, s)[mi] : T(0); };
    const int pi = i - 1 - c.sub, pj = j - 1 - c.sub;  // profile index
    if (fw_in(c, 1, i, j)) {
        const T lb = c.sub ? lam[ij] : T(0.5) * (lam[ij + P] + lam[ij]);
        const T mb = c.sub ? mu[ij] : T(0.5) * (mu[ij + P] + mu[ij]);
        const T lm = lb + T(2) * mb;
        T dx = (T(27) * vx[ij + P] - T(27) * vx[ij] - vx[ij + 2 * P] + vx[ij - P]) * c.ihx;
        T dy = (T(27) * vy[ij] - T(27) * vy[ij - 1] - vy[ij + 1] + vy[ij - 2]) * c.ihy;
        m1 = c.bxh[pi] * mold(M1) + c.axh[pi] * dx;
        m2 = c.byf[pj] * mold(M2) + c.ayf[pj] * dy;
        dx += m1;
        dy += m2;
        sxx += (lm * dx + lb * dy) * c.dt;
        syy += (lm * dy + lb * dx) * c.dt;
    }
    if (fw_in(c, 2, i, j)) {
        const T mb = c.sub ? mu[ij] : T(0.5) * (mu[ij] + mu[ij + 1]);
        T dx = (T(27) * vy[ij] - T(27) * vy[ij - P] - vy[ij + P] + vy[ij - 2 * P]) * c.ihx;
        T dy = (T(27) * vx[ij + 1] - T(27) * vx[ij] - vx[ij + 2] + vx[ij - 1]) * c.ihy;
        m3 = c.bxf[pi] * mold(M3) + c.axf[pi] * dx;
        m4 = c.byh[pj] * mold(M4) + c.ayh[pj] * dy;
        sxy += mb * (dx + m3 + dy + m4) * c.dt;
    }
    O.f(SXX, s)[ij] = sxx;
    O.f(SYY, s)[ij] = syy;
    O.f(SXY, s)[ij] = sxy;
    if (mo >= 0) {
        O.f(M1, s)[mo] = m1;
        O.f(M2, s)[mo] = m2;
        O.f(M3, s)[mo] = m3;
        O.f(M4, s)[mo] = m4;
    }
}

// one cell of fw3/fw4 (reads the new stresses of level n)
template <typename T>
__device__ inline void el_vel_cell(const ElC<T> &I, const El<T> &O, const T *__restrict__ rho,
                                   const ElCoef<T> &c, int s, int i, int j)
{
    const int P = c.P;
    const int64_t ij = (int64_t)i * P + j;
    const T *__restrict__ sxx = O.f(SXX, s);
    const T *__restrict__ syy = O.f(SYY, s);
    const T *__restrict__ sxy = O.f(SXY, s);
    T vx = I.f(VX, s)[ij], vy = I.f(VY, s)[ij];
    T m5 = 0, m6 = 0, m7 = 0, m8 = 0;
    const int64_t mi = I.mv(i, j, ij), mo = O.mv(i, j, ij);  // memory variables (< 0: all 0)
    auto mold = [&](int q) { return mi >= 0 ? I.f(q, s)[mi] : T(0); };
    const int pi = i - 1 - c.sub, pj = j - 1 - c.sub;  // profile index
    if (fw_in(c, 3, i, j)) {
        T dx = (T(27) * sxx[ij] - T(27) * sxx[ij - P] - sxx[ij + P] + sxx[ij - 2 * P]) * c.ihx;
        T dy = (T(27) * sxy[ij] - T(27) * sxy[ij - 1] - sxy[ij + 1] + sxy[ij - 2]) * c.ihy;
        m5 = c.bxf[pi] * mold(M5) + c.axf[pi] * dx;
        m6 = c.byf[pj] * mold(M6) + c.ayf[pj] * dy;
        vx += (dx + m5 + dy + m6) * c.dt / rho[ij];
    }
    if (fw_in(c, 4, i, j)) {
        const T rb = c.sub ? rho[ij] : T(0.25) * (rho[ij] + rho[ij + P] + rho[ij + P + 1] + rho[ij + 1]);
        T dx = (T(27) * sxy[ij + P] - T(27) * sxy[ij] - sxy[ij + 2 * P] + sxy[ij - P]) * c.ihx;
        T dy = (T(27) * syy[ij + 1] - T(27) * syy[ij] - syy[ij + 2] + syy[ij - 1]) * c.ihy;
        m7 = c.bxh[pi] * mold(M7) + c.axh[pi] * dx;
        m8 = c.byh[pj] * mold(M8) + c.ayh[pj] * dy;
        vy += (dx + m7 + dy + m8) * c.dt / rb;
    }
    O.f(VX, s)[ij] = vx;
    O.f(VY, s)[ij] = vy;
    if (mo >= 0) {
        O.f(M5, s)[mo] = m5;
        O.f(M6, s)[mo] = m6;
        O.f(M7, s)[mo] = m7;
        O.f(M8, s)[mo] = m8;
    }
}

// frame: one column per lane, FRPT rows per wavefront item, one shot per grid z
template <typename T>
__global__ __launch_bounds__(WAVE *NWAVES) void el_stress_frame(ElC<T> I, El<T> O, const T *__restrict__ lam,
                                                                const T *__restrict__ mu, ElCoef<T> c,
                                                                const int4 *__restrict__ items, int nitems)
{
    const int it = blockIdx.x * NWAVES + wave_id();
    if (it >= nitems) return;
    const int4 m = items[it];
    const int j = m.z + (threadIdx.x & (WAVE - 1));
    if (j >= m.w) return;
    for (int r = 0; r < m.y; r++) el_stress_cell(I, O, lam, mu, c, blockIdx.z, m.x + r, j);
}

template <typename T>
__global__ __launch_bounds__(WAVE *NWAVES) void el_vel_frame(ElC<T> I, El<T> O, const T *__restrict__ rho,
                                                             ElCoef<T> c, const int4 *__restrict__ items,
                                                             int nitems)
{
    const int it = blockIdx.x * NWAVES + wave_id();
    if (it >= nitems) return;
    const int4 m = items[it];
    const int j = m.z + (threadIdx.x & (WAVE - 1));
    if (j >= m.w) return;
    for (int r = 0; r < m.y; r++) el_vel_cell(I, O, rho, c, blockIdx.z, m.x + r, j);
}

// Row neighbours at column offsets -2, -1, VEC, VEC+1 of a lane's vector: from the lanes that
// own them or, where that lane is outside [0, last], from the lane's own load (issued with the
// other loads of the shot).
template <int VEC> __device__ constexpr int hoff(int q) { return q < 2 ? q - 2 : VEC + q - 2; }
template <int VEC> __device__ constexpr int hlane(int q)
{
    return hoff<VEC>(q) >= 0 ? hoff<VEC>(q) / VEC : -((-hoff<VEC>(q) + VEC - 1) / VEC);
}
template <typename T, int VEC>
__device__ inline void halo_load(T (&h)[4], const T *__restrict__ row, int j0, int lane, int last, bool ok)
{
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int src = lane + hlane<VEC>(q);
        h[q] = ok && (src < 0 || src > last) ? row[j0 + hoff<VEC>(q)] : T(0);
    }
}
// ext[k + 2] = value at column j0 + k, k = -2 .. VEC+1
template <typename T, int VEC>
__device__ inline void halo_fill(T (&ext)[VEC + 4], const T (&x)[VEC], const T (&h)[4], int lane, int last)
{
#pragma unroll
    for (int e = 0; e < VEC; e++) ext[e + 2] = x[e];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int d = hlane<VEC>(q);
        const int el = hoff<VEC>(q) - d * VEC;
        const T v = d < 0 ? __shfl_up(x[el], -d) : __shfl_down(x[el], d);
        const int src = lane + d;
        ext[hoff<VEC>(q) + 2] = (src < 0 || src > last) ? h[q] : v;
    }
}

// Rectangle stress update (fw1 + fw2 without memory variables).  A wavefront owns RPT rows x
// (64 lanes x VEC depth cells); a workgroup loops over a group of Sg shots with lambda and mu
// in registers.
template <typename T, int RPT, int VEC>
__device__ inline void el_stress_rect_body(const ElC<T> &I, const El<T> &O, const T *__restrict__ lam,
                                           const T *__restrict__ mu, const ElCoef<T> &c, const Rect &R,
                                           const TileMap &map, int act, int Sg, int bid)
{
    int g, strip, tj;
    if (!decode_tile(map, g, strip, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int jb = R.jL + tj * WAVE * VEC, j0 = jb + lane * VEC;
    const int ib = R.iL + (strip * NWAVES + wv) * RPT;
    if (ib >= R.iR) return;  // wave-uniform
    const int P = c.P, nr = min(RPT, R.iR - ib);
    const bool on = j0 < R.jR;
    const int last = min(WAVE - 1, (R.jR - jb) / VEC - 1);
    T La[RPT + 1][VEC], Mu[RPT + 1][VEC], MuE[RPT];
#pragma unroll
    for (int q = 0; q <= RPT; q++) {
        if (on && q <= nr) {
            vld(La[q], lam + (int64_t)(ib + q) * P + j0);
            vld(Mu[q], mu + (int64_t)(ib + q) * P + j0);
        } else {
            vzero(La[q]);
            vzero(Mu[q]);
        }
    }
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        T h[4], ext[VEC + 4];
        halo_load<T, VEC>(h, mu + (int64_t)(ib + r) * P, j0, lane, last, on && r < nr);
        halo_fill<T, VEC>(ext, Mu[r], h, lane, last);
        MuE[r] = ext[VEC + 2];
    }
    const int s1 = min(act, (g + 1) * Sg);
    for (int s = g * Sg; s < s1; s++) {
        const T *__restrict__ vx = I.f(VX, s);
        const T *__restrict__ vy = I.f(VY, s);
        T X[RPT + 3][VEC], Y[RPT + 3][VEC], HX[RPT][4], HY[RPT][4];
        T S0[RPT][VEC], S1[RPT][VEC], S2[RPT][VEC];
#pragma unroll
        for (int q = 0; q < RPT + 3; q++) {  // vx rows ib-1 .., vy rows ib-2 .. (inside [0, NX+1])
            if (on && q <= nr + 2) {
                vld(X[q], vx + (int64_t)(ib - 1 + q) * P + j0);
                vld(Y[q], vy + (int64_t)(ib - 2 + q) * P + j0);
            } else {
                vzero(X[q]);
                vzero(Y[q]);
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int64_t row = (int64_t)(ib + r) * P;
            const bool ok = on && r < nr;
            halo_load<T, VEC>(HX[r], vx + row, j0, lane, last, ok);
            halo_load<T, VEC>(HY[r], vy + row, j0, lane, last, ok);
            if (ok) {
                vld(S0[r], I.f(SXX, s) + row + j0);
                vld(S1[r], I.f(SYY, s) + row + j0);
                vld(S2[r], I.f(SXY, s) + row + j0);
            } else {
                vzero(S0[r]);
                vzero(S1[r]);
                vzero(S2[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (r >= nr) break;  // uniform
            T xe[VEC + 4], ye[VEC + 4];
            halo_fill<T, VEC>(xe, X[r + 1], HX[r], lane, last);
            halo_fill<T, VEC>(ye, Y[r + 2], HY[r], lane, last);
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const T lb = c.sub ? La[r][e] : T(0.5) * (La[r + 1][e] + La[r][e]);
                const T mb = c.sub ? Mu[r][e] : T(0.5) * (Mu[r + 1][e] + Mu[r][e]);
                const T lm = lb + T(2) * mb;
                const T dx = (T(27) * X[r + 2][e] - T(27) * X[r + 1][e] - X[r + 3][e] + X[r][e]) * c.ihx;
                const T dy = (T(27) * ye[e + 2] - T(27) * ye[e + 1] - ye[e + 3] + ye[e]) * c.ihy;
                S0[r][e] += (lm * dx + lb * dy) * c.dt;
                S1[r][e] += (lm * dy + lb * dx) * c.dt;
                const T mb2 = c.sub ? Mu[r][e]
                                    : T(0.5) * (Mu[r][e] + (e + 1 < VEC ? Mu[r][e + 1 < VEC ? e + 1 : 0] : MuE[r]));
                const T dx2 = (T(27) * Y[r + 2][e] - T(27) * Y[r + 1][e] - Y[r + 3][e] + Y[r][e]) * c.ihx;
                const T dy2 = (T(27) * xe[e + 3] - T(27) * xe[e + 2] - xe[e + 4] + xe[e + 1]) * c.ihy;
                S2[r][e] += mb2 * (dx2 + dy2) * c.dt;
            }
            if (on) {
                const int64_t row = (int64_t)(ib + r) * P;
                vst(O.f(SXX, s) + row + j0, S0[r]);
                vst(O.f(SYY, s) + row + j0, S1[r]);
                vst(O.f(SXY, s) + row + j0, S2[r]);
            }
        }
    }
}

// Rectangle velocity update (fw3 + fw4 without memory variables), reading the new stresses.
template <typename T, int RPT, int VEC>
__device__ inline void el_vel_rect_body(const ElC<T> &I, const El<T> &O, const T *__restrict__ rho,
                                        const ElCoef<T> &c, const Rect &R, const TileMap &map, int act,
                                        int Sg, int bid)
{
    int g, strip, tj;
    if (!decode_tile(map, g, strip, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int jb = R.jL + tj * WAVE * VEC, j0 = jb + lane * VEC;
    const int ib = R.iL + (strip * NWAVES + wv) * RPT;
    if (ib >= R.iR) return;  // wave-uniform
    const int P = c.P, nr = min(RPT, R.iR - ib);
    const bool on = j0 < R.jR;
    const int last = min(WAVE - 1, (R.jR - jb) / VEC - 1);
    T Rh[RPT + 1][VEC], RhE[RPT + 1];
#pragma unroll
    for (int q = 0; q <= RPT; q++) {
        if (on && q <= nr) vld(Rh[q], rho + (int64_t)(ib + q) * P + j0);
        else vzero(Rh[q]);
        T h[4], ext[VEC + 4];
        halo_load<T, VEC>(h, rho + (int64_t)(ib + q) * P, j0, lane, last, on && q <= nr);
        halo_fill<T, VEC>(ext, Rh[q], h, lane, last);
        RhE[q] = ext[VEC + 2];
    }
    const int s1 = min(act, (g + 1) * Sg);
    for (int s = g * Sg; s < s1; s++) {
        const T *__restrict__ sxx = O.f(SXX, s);
        const T *__restrict__ syy = O.f(SYY, s);
        const T *__restrict__ sxy = O.f(SXY, s);
        T SX[RPT + 3][VEC], SY[RPT + 3][VEC], YY[RPT][VEC], HS[RPT][4], HY[RPT][4];
        T VXr[RPT][VEC], VYr[RPT][VEC];
#pragma unroll
        for (int q = 0; q < RPT + 3; q++) {  // sxx rows ib-2 .., sxy rows ib-1 ..
            if (on && q <= nr + 2) {
                vld(SX[q], sxx + (int64_t)(ib - 2 + q) * P + j0);
                vld(SY[q], sxy + (int64_t)(ib - 1 + q) * P + j0);
            } else {
                vzero(SX[q]);
                vzero(SY[q]);
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int64_t row = (int64_t)(ib + r) * P;
            const bool ok = on && r < nr;
            halo_load<T, VEC>(HS[r], sxy + row, j0, lane, last, ok);
            halo_load<T, VEC>(HY[r], syy + row, j0, lane, last, ok);
            if (ok) {
                vld(YY[r], syy + row + j0);
                vld(VXr[r], I.f(VX, s) + row + j0);
                vld(VYr[r], I.f(VY, s) + row + j0);
            } else {
                vzero(YY[r]);
                vzero(VXr[r]);
                vzero(VYr[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (r >= nr) break;  // uniform
            T se[VEC + 4], ye[VEC + 4];
            halo_fill<T, VEC>(se, SY[r + 1], HS[r], lane, last);
            halo_fill<T, VEC>(ye, YY[r], HY[r], lane, last);
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const T dx = (T(27) * SX[r + 2][e] - T(27) * SX[r + 1][e] - SX[r + 3][e] + SX[r][e]) * c.ihx;
                const T dy = (T(27) * se[e + 2] - T(27) * se[e + 1] - se[e + 3] + se[e]) * c.ihy;
                VXr[r][e] += (dx + dy) * c.dt / Rh[r][e];
                const T re0 = e + 1 < VEC ? Rh[r][e + 1 < VEC ? e + 1 : 0] : RhE[r];
                const T re1 = e + 1 < VEC ? Rh[r + 1][e + 1 < VEC ? e + 1 : 0] : RhE[r + 1];
                const T rb = c.sub ? Rh[r][e] : T(0.25) * (Rh[r][e] + Rh[r + 1][e] + re1 + re0);
                const T dx2 = (T(27) * SY[r + 2][e] - T(27) * SY[r + 1][e] - SY[r + 3][e] + SY[r][e]) * c.ihx;
                const T dy2 = (T(27) * ye[e + 3] - T(27) * ye[e + 2] - ye[e + 4] + ye[e + 1]) * c.ihy;
                VYr[r][e] += (dx2 + dy2) * c.dt / rb;
            }
            if (on) {
                const int64_t row = (int64_t)(ib + r) * P;
                vst(O.f(VX, s) + row + j0, VXr[r]);
                vst(O.f(VY, s) + row + j0, VYr[r]);
            }
        }
    }
}

// Frame work in streaming form (the fused forward launches).  An item is nr rows of at most
// SW_FCOLS columns [m.z, m.w) outside the rectangle; lane L holds column m.z - 2 + L, so two
// lanes on each side carry the columns the y-direction stencils reach.  For its shot the
// wavefront walks the item's rows, holding the x-direction stencil inputs of the rows around
// the current one in registers (one new row loaded per step) and taking the y-direction
// neighbours from the adjacent lanes; each cell is el_stress_cell / el_vel_cell, the same
// expressions, with its memory variables at their index in the level's layout.
// columns / rows per streaming frame item (measured at 600^2 x 16 shots, forward level: 1, 2,
// 4, 8, 16, 32 rows 128.2, 118.8, 115.7, 113.2, 121.4, 169.2 us: taller items leave too few
// wavefronts, shorter ones re-read their halo rows)
constexpr int SW_FCOLS = WAVE - 4, FRS_ROWS = 8;
template <typename T> __device__ inline T fsh(T v, int src) { return __shfl(v, src, WAVE); }

template <typename T>
__device__ inline void stress_frame_strm(const ElC<T> &I, const El<T> &O, const T *__restrict__ lam,
                                         const T *__restrict__ mu, const ElCoef<T> &c, int s, int4 m)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int j = m.z - 2 + lane;
    const int P = c.P, NXp = c.NX + 1, NYp = c.NY + 1;
    const int jc = j < 0 ? 0 : (j > NYp ? NYp : j);  // loads of lanes beyond the grid stay inside it
    const bool own = lane >= 2 && j < m.w;
    const int pj = j - 1 - c.sub;
    const T *__restrict__ vx = I.f(VX, s);
    const T *__restrict__ vy = I.f(VY, s);
    // row r of an array (0 outside the grid; such values only reach cells outside fw1 / fw2)
    auto rowv = [&](const T *__restrict__ f, int r) { return r >= 0 && r <= NXp ? f[(int64_t)r * P + jc] : T(0); };
    const int i0 = m.x, i1 = m.x + m.y;
    // windows: vx rows x-1 .. x+2 (X0 .. X3), vy rows x-2 .. x+1 (Y0 .. Y3), lam and mu rows x, x+1
    T X0 = rowv(vx, i0 - 1), X1 = rowv(vx, i0), X2 = rowv(vx, i0 + 1), X3;
    T Y0 = rowv(vy, i0 - 2), Y1 = rowv(vy, i0 - 1), Y2 = rowv(vy, i0), Y3;
    T L0 = rowv(lam, i0), L1, U0 = rowv(mu, i0), U1;
    for (int i = i0; i < i1; i++) {
        X3 = rowv(vx, i + 2);
        Y3 = rowv(vy, i + 1);
        L1 = rowv(lam, i + 1);
        U1 = rowv(mu, i + 1);
        // row i's y-direction neighbours: vy at j-2 .. j+1, vx at j-1 .. j+2, mu at j+1
        const T vym = fsh(Y2, lane - 1), vyp = fsh(Y2, lane + 1), vymm = fsh(Y2, lane - 2);
        const T vxm = fsh(X1, lane - 1), vxp = fsh(X1, lane + 1), vxpp = fsh(X1, lane + 2);
        const T mue = fsh(U0, lane + 1);
        if (own) {
            const int64_t ij = (int64_t)i * P + j;
            T sxx = I.f(SXX, s)[ij], syy = I.f(SYY, s)[ij], sxy = I.f(SXY, s)[ij];
            T m1 = 0, m2 = 0, m3 = 0, m4 = 0;
            const int64_t mi = I.mv(i, j, ij), mo = O.mv(i, j, ij);  // memory variables (< 0: all 0)
            auto mold = [&](int q) { return mi >= 0 ? I.f(q, s)[mi] : T(0); };
            const int pi = i - 1 - c.sub;  // profile index
            if (fw_in(c, 1, i, j)) {
                const T lb = c.sub ? L0 : T(0.5) * (L1 + L0);
                const T mb = c.sub ? U0 : T(0.5) * (U1 + U0);
                const T lm = lb + T(2) * mb;
                T dx = (T(27) * X2 - T(27) * X1 - X3 + X0) * c.ihx;
                T dy = (T(27) * Y2 - T(27) * vym - vyp + vymm) * c.ihy;
                m1 = c.bxh[pi] * mold(M1) + c.axh[pi] * dx;
                m2 = c.byf[pj] * mold(M2) + c.ayf[pj] * dy;
                dx += m1;
                dy += m2;
                sxx += (lm * dx + lb * dy) * c.dt;
                syy += (lm * dy + lb * dx) * c.dt;
            }
            if (fw_in(c, 2, i, j)) {
                const T mb = c.sub ? U0 : T(0.5) * (U0 + mue);
                T dx = (T(27) * Y2 - T(27) * Y1 - Y3 + Y0) * c.ihx;
                T dy = (T(27) * vxp - T(27) * X1 - vxpp + vxm) * c.ihy;
                m3 = c.bxf[pi] * mold(M3) + c.axf[pi] * dx;
                m4 = c.byh[pj] * mold(M4) + c.ayh[pj] * dy;
                sxy += mb * (dx + m3 + dy + m4) * c.dt;
            }
            O.f(SXX, s)[ij] = sxx;
            O.f(SYY, s)[ij] = syy;
            O.f(SXY, s)[ij] = sxy;
            if (mo >= 0) {
                O.f(M1, s)[mo] = m1;
                O.f(M2, s)[mo] = m2;
                O.f(M3, s)[mo] = m3;
                O.f(M4, s)[mo] = m4;
            }
        }
        X0 = X1;
        X1 = X2;
        X2 = X3;
        Y0 = Y1;
        Y1 = Y2;
        Y2 = Y3;
        L0 = L1;
        U0 = U1;
    }
}

template <typename T>
__device__ inline void vel_frame_strm(const ElC<T> &I, const El<T> &O, const T *__restrict__ rho,
                                      const ElCoef<T> &c, int s, int4 m)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int j = m.z - 2 + lane;
    const int P = c.P, NXp = c.NX + 1, NYp = c.NY + 1;
    const int jc = j < 0 ? 0 : (j > NYp ? NYp : j);
    const bool own = lane >= 2 && j < m.w;
    const int pj = j - 1 - c.sub;
    const T *__restrict__ sxx = O.f(SXX, s);
    const T *__restrict__ syy = O.f(SYY, s);
    const T *__restrict__ sxy = O.f(SXY, s);
    auto rowv = [&](const T *__restrict__ f, int r) { return r >= 0 && r <= NXp ? f[(int64_t)r * P + jc] : T(0); };
    const int i0 = m.x, i1 = m.x + m.y;
    // windows: sxx rows x-2 .. x+1 (A0 .. A3), sxy rows x-1 .. x+2 (B0 .. B3), rho rows x, x+1
    T A0 = rowv(sxx, i0 - 2), A1 = rowv(sxx, i0 - 1), A2 = rowv(sxx, i0), A3;
    T B0 = rowv(sxy, i0 - 1), B1 = rowv(sxy, i0), B2 = rowv(sxy, i0 + 1), B3;
    T R0 = rowv(rho, i0), R1;
    for (int i = i0; i < i1; i++) {
        A3 = rowv(sxx, i + 1);
        B3 = rowv(sxy, i + 2);
        R1 = rowv(rho, i + 1);
        const T YY = rowv(syy, i);
        // row i's y-direction neighbours: sxy at j-2 .. j+1, syy at j-1 .. j+2, rho at j+1
        const T xym = fsh(B1, lane - 1), xyp = fsh(B1, lane + 1), xymm = fsh(B1, lane - 2);
        const T yym = fsh(YY, lane - 1), yyp = fsh(YY, lane + 1), yypp = fsh(YY, lane + 2);
        const T re0 = fsh(R0, lane + 1), re1 = fsh(R1, lane + 1);
        if (own) {
            const int64_t ij = (int64_t)i * P + j;
            T vxn = I.f(VX, s)[ij], vyn = I.f(VY, s)[ij];
            T m5 = 0, m6 = 0, m7 = 0, m8 = 0;
            const int64_t mi = I.mv(i, j, ij), mo = O.mv(i, j, ij);  // memory variables (< 0: all 0)
            auto mold = [&](int q) { return mi >= 0 ? I.f(q, s)[mi] : T(0); };
            const int pi = i - 1 - c.sub;  // profile index
            if (fw_in(c, 3, i, j)) {
                T dx = (T(27) * A2 - T(27) * A1 - A3 + A0) * c.ihx;
                T dy = (T(27) * B1 - T(27) * xym - xyp + xymm) * c.ihy;
                m5 = c.bxf[pi] * mold(M5) + c.axf[pi] * dx;
                m6 = c.byf[pj] * mold(M6) + c.ayf[pj] * dy;
                vxn += (dx + m5 + dy + m6) * c.dt / R0;
            }
            if (fw_in(c, 4, i, j)) {
                const T rb = c.sub ? R0 : T(0.25) * (R0 + R1 + re1 + re0);
                T dx = (T(27) * B2 - T(27) * B1 - B3 + B0) * c.ihx;
                T dy = (T(27) * yyp - T(27) * YY - yypp + yym) * c.ihy;
                m7 = c.bxh[pi] * mold(M7) + c.axh[pi] * dx;
                m8 = c.byh[pj] * mold(M8) + c.ayh[pj] * dy;
                vyn += (dx + m7 + dy + m8) * c.dt / rb;
            }
            O.f(VX, s)[ij] = vxn;
            O.f(VY, s)[ij] = vyn;
            if (mo >= 0) {
                O.f(M5, s)[mo] = m5;
                O.f(M6, s)[mo] = m6;
                O.f(M7, s)[mo] = m7;
                O.f(M8, s)[mo] = m8;
            }
        }
        A0 = A1;
        A1 = A2;
        A2 = A3;
        B0 = B1;
        B1 = B2;
        B2 = B3;
        R0 = R1;
    }
}

// Kernels: the rectangle alone, and the rectangle with the frame in one launch (pitched
// layout): workgroups [0, nfb) take the frame items (nib item blocks of NWAVES items per shot),
// the rest the rectangle tiles, so the latency-bound frame work overlaps the rectangle's.
template <typename T, int RPT, int VEC>
__global__ __launch_bounds__(WAVE *NWAVES) void el_stress_rect(ElC<T> I, El<T> O, const T *__restrict__ lam,
                                                               const T *__restrict__ mu, ElCoef<T> c,
                                                               Rect R, TileMap map, int act, int Sg)
{
    el_stress_rect_body<T, RPT, VEC>(I, O, lam, mu, c, R, map, act, Sg, (int)blockIdx.x);
}
template <typename T, int RPT, int VEC>
__global__ __launch_bounds__(WAVE *NWAVES) void el_vel_rect(ElC<T> I, El<T> O, const T *__restrict__ rho,
                                                            ElCoef<T> c, Rect R, TileMap map, int act,
                                                            int Sg)
{
    el_vel_rect_body<T, RPT, VEC>(I, O, rho, c, R, map, act, Sg, (int)blockIdx.x);
}
template <typename T, int RPT, int VEC>
__global__ __launch_bounds__(WAVE *NWAVES) void el_stress_level(ElC<T> I, El<T> O, const T *__restrict__ lam,
                                                                const T *__restrict__ mu, ElCoef<T> c,
                                                                Rect R, TileMap map, int act, int Sg,
                                                                const int4 *__restrict__ items, int nitems,
                                                                int nib, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b >= nfb) {
        el_stress_rect_body<T, RPT, VEC>(I, O, lam, mu, c, R, map, act, Sg, b - nfb);
        return;
    }
    const int it = (b % nib) * NWAVES + wave_id(), s = b / nib;
    if (it >= nitems || s >= act) return;  // wave-uniform
    stress_frame_strm(I, O, lam, mu, c, s, items[it]);
}
template <typename T, int RPT, int VEC>
__global__ __launch_bounds__(WAVE *NWAVES) void el_vel_level(ElC<T> I, El<T> O, const T *__restrict__ rho,
                                                             ElCoef<T> c, Rect R, TileMap map, int act,
                                                             int Sg, const int4 *__restrict__ items,
                                                             int nitems, int nib, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b >= nfb) {
        el_vel_rect_body<T, RPT, VEC>(I, O, rho, c, R, map, act, Sg, b - nfb);
        return;
    }
    const int it = (b % nib) * NWAVES + wave_id(), s = b / nib;
    if (it >= nitems || s >= act) return;  // wave-uniform
    vel_frame_strm(I, O, rho, c, s, items[it]);
}

// Sources at level n (unique (cell, type) keys, entries in source order), then receivers:
// trace[r * (nstep_s + 1) + n] = field_type_r[cell_r] (GetReceive.cpp:13-45 layout).
// mode 0: add; 1: add, then record the receivers; 2: put back the stresses of level n as
// fw3/fw4 read them, before injection (the adjoint needs those).  pre (one value per key,
// may be null) keeps the stress at each stress-type key before the add (modes 0 and 1) and
// gives it back in mode 2, so the adjoint sees the exact pre-injection value; without it,
// mode 2 subtracts the sources again, which rounds in FP32.
template <typename T>
__device__ inline void el_src_rcv_body(
    const El<T> &O, int n, const int64_t *__restrict__ us_ptr, const int64_t *__restrict__ us_key,
    const int64_t *__restrict__ us_eptr, const int64_t *__restrict__ us_ent,
    const T *__restrict__ srcv, const int64_t *__restrict__ srcv_off,
    const int64_t *__restrict__ nsrc, const int64_t *__restrict__ rcv_ptr,
    const int64_t *__restrict__ rcv_key, const int64_t *__restrict__ nstep,
    T *__restrict__ traces, const int64_t *__restrict__ trace_off, int mode, T *__restrict__ pre, int s)
{
    const int64_t ns = nsrc[s];
    const T *__restrict__ row = srcv + srcv_off[s] + (int64_t)(n - 1) * ns;
    for (int64_t c = us_ptr[s] + threadIdx.x; c < us_ptr[s + 1]; c += blockDim.x) {
        const int64_t key = us_key[c];
        const int type = (int)(key & 7);
        const int64_t cell = key >> 3;
        if (mode == 2 && type < SXX) continue;  // removal touches stresses only
        T *__restrict__ f = O.f(type, s);
        if (mode == 2 && pre) {
            f[cell] = pre[c];
            continue;
        }
        T acc = T(0);
        for (int64_t e = us_eptr[c]; e < us_eptr[c + 1]; e++) acc += row[us_ent[e]];
        if (mode != 2 && pre && type >= SXX) pre[c] = f[cell];
        f[cell] = mode == 2 ? f[cell] - acc : f[cell] + acc;
    }
    if (mode != 1) return;
    __syncthreads();
    const int64_t r0 = rcv_ptr[s], nr = rcv_ptr[s + 1] - r0, nt = nstep[s] + 1;
    T *__restrict__ tr = traces + trace_off[s];
    // RB receivers per thread and pass: every key load, then every field load, then the
    // stores, so a pass waits on two load latencies instead of two per receiver
    constexpr int RB = 8;
    for (int64_t rb = threadIdx.x; rb < nr; rb += (int64_t)RB * blockDim.x) {
        int64_t key[RB];
        T v[RB];
#pragma unroll
        for (int k = 0; k < RB; k++) {
            const int64_t r = rb + (int64_t)k * blockDim.x;
            key[k] = r < nr ? rcv_key[r0 + r] : -1;
        }
#pragma unroll
        for (int k = 0; k < RB; k++) v[k] = key[k] >= 0 ? O.f((int)(key[k] & 7), s)[key[k] >> 3] : T(0);
#pragma unroll
        for (int k = 0; k < RB; k++) {
            const int64_t r = rb + (int64_t)k * blockDim.x;
            if (r < nr) tr[r * nt + n] = v[k];
        }
    }
}
template <typename T>
__global__ __launch_bounds__(256) void el_src_rcv(
    El<T> O, int n, const int64_t *__restrict__ us_ptr, const int64_t *__restrict__ us_key,
    const int64_t *__restrict__ us_eptr, const int64_t *__restrict__ us_ent,
    const T *__restrict__ srcv, const int64_t *__restrict__ srcv_off,
    const int64_t *__restrict__ nsrc, const int64_t *__restrict__ rcv_ptr,
    const int64_t *__restrict__ rcv_key, const int64_t *__restrict__ nstep,
    T *__restrict__ traces, const int64_t *__restrict__ trace_off, int mode, T *__restrict__ pre)
{
    el_src_rcv_body(O, n, us_ptr, us_key, us_eptr, us_ent, srcv, srcv_off, nsrc, rcv_ptr, rcv_key, nstep,
                    traces, trace_off, mode, pre, (int)blockIdx.x);
}

// ------------------------------- adjoint -------------------------------------------------
// Per level, each phase (velocity, then stress) runs three kernels over the forward's split:
//   frame_a : on frame cells, the cotangents t0..t3 of the four derivative terms (to Tb),
//             including the CPML memory recursion, and the parameter terms;
//   rect    : on rectangle cells, the parameter terms and the transposed-stencil gathers into
//             the stress (velocity) cotangents.  There the CPML is the identity, so t0..t3
//             are recomputed from the field cotangents (t0 = t1, t2 = t3) instead of stored;
//   frame_b : the same gathers on frame cells.
// Parameter terms are accumulated per shot group at the cell where the forward formula uses
// the averaged parameter (rho at fw3, rho-bar at fw4, lambda-bar and mu-bar at fw1, mu-bar at
// fw2) and mapped onto (rho, lambda, mu) once after the sweep (el_grad_map).
enum { G_RHO = 0, G_RB, G_LB, G_MB1, G_MB2, G_COUNT };

template <typename T> struct ElAdj {
    El<T> A;          // cotangents of the 13 state arrays
    ElC<T> Sn, Sm;    // forward state of levels n (this level's sources removed) and n-1
    T *Tb;            // [4][shot][Npad]: t0..t3 on frame cells
    T *mo[8];         // full-grid kernels: where M1..M8 cotangents of level n-1 go ([shot][Npad])
    T *G;             // [G_COUNT][group][Npad]
    int64_t npad, tstride, gstride;  // Npad, nshots * Npad, ngroups * Npad
    const T *rho, *lam, *mu, *rb;    // rb: rho-bar of fw4 (Core.jl:201)
    Rect R;
    int act, Sg;
};

template <typename T> __device__ inline bool in_rect(const ElAdj<T> &a, int i, int j)
{
    return i >= a.R.iL && i < a.R.iR && j >= a.R.jL && j < a.R.jR;
}
template <typename T> __device__ inline T tb_at(const ElAdj<T> &a, int k, int s, int64_t ij)
{
    return a.Tb[(int64_t)k * a.tstride + (int64_t)s * a.npad + ij];
}
template <typename T> __device__ inline T &g_at(const ElAdj<T> &a, int k, int g, int64_t ij)
{
    return a.G[(int64_t)k * a.gstride + (int64_t)g * a.npad + ij];
}

// t_k of the velocity phase at (i, j), shot s; 0 outside the interior (Core.jl:215-228 ranges)
template <typename T>
__device__ inline T tv_vel(const ElAdj<T> &a, const ElCoef<T> &c, int k, int s, int i, int j)
{
    if (i < 1 || i > c.NX || j < 1 || j > c.NY) return T(0);
    const int64_t ij = (int64_t)i * c.P + j;
    if (in_rect(a, i, j))
        return k < 2 ? a.A.f(VX, s)[ij] * c.dt / a.rho[ij] : a.A.f(VY, s)[ij] * c.dt / a.rb[ij];
    return tb_at(a, k, s, ij);
}
// t_k of the stress phase
template <typename T>
__device__ inline T tv_str(const ElAdj<T> &a, const ElCoef<T> &c, int k, int s, int i, int j)
{
    if (i < 1 || i > c.NX || j < 1 || j > c.NY) return T(0);
    const int64_t ij = (int64_t)i * c.P + j;
    if (in_rect(a, i, j)) {
        const int P = c.P;
        if (k < 2) {
            const T lb = c.sub ? a.lam[ij] : T(0.5) * (a.lam[ij + P] + a.lam[ij]);
            const T mb = c.sub ? a.mu[ij] : T(0.5) * (a.mu[ij + P] + a.mu[ij]);
            const T lm = lb + T(2) * mb;
            const T gx = a.A.f(SXX, s)[ij] * c.dt, gy = a.A.f(SYY, s)[ij] * c.dt;
            return k == 0 ? lm * gx + lb * gy : lb * gx + lm * gy;
        }
        const T mb = c.sub ? a.mu[ij] : T(0.5) * (a.mu[ij] + a.mu[ij + 1]);
        return a.A.f(SXY, s)[ij] * mb * c.dt;
    }
    return tb_at(a, k, s, ij);
}

// velocity phase, frame_a: fw3/fw4 transposed through the memory recursion (m5..m8)
template <typename T>
__device__ inline void adjv_a_cell(const ElAdj<T> &a, const ElCoef<T> &c, int s, int i, int j,
                                   T &gR, T &gB)
{
    const int P = c.P;
    const int64_t ij = (int64_t)i * P + j;
    T t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    T *a5 = a.A.f(M5, s), *a6 = a.A.f(M6, s), *a7 = a.A.f(M7, s), *a8 = a.A.f(M8, s);
    T n5 = 0, n6 = 0, n7 = 0, n8 = 0;
    const T *__restrict__ sxx = a.Sn.f(SXX, s);
    const T *__restrict__ syy = a.Sn.f(SYY, s);
    const T *__restrict__ sxy = a.Sn.f(SXY, s);
    const int64_t mi = a.Sn.mv(i, j, ij);  // forward memory variables of level n (< 0: all 0)
    auto msn = [&](int q) { return mi >= 0 ? a.Sn.f(q, s)[mi] : T(0); };
    const int pi = i - 1 - c.sub, pj = j - 1 - c.sub;  // profile index
    if (fw_in(c, 3, i, j)) {
        const T dx = (T(27) * sxx[ij] - T(27) * sxx[ij - P] - sxx[ij + P] + sxx[ij - 2 * P]) * c.ihx;
        const T dy = (T(27) * sxy[ij] - T(27) * sxy[ij - 1] - sxy[ij + 1] + sxy[ij - 2]) * c.ihy;
        const T S3 = dx + msn(M5) + dy + msn(M6);
        const T g = a.A.f(VX, s)[ij], r = a.rho[ij];
        gR += -g * S3 * c.dt / (r * r);
        const T e = g * c.dt / r;
        const T g5 = a5[ij] + e, g6 = a6[ij] + e;
        t0 = e + c.axf[pi] * g5;
        t1 = e + c.ayf[pj] * g6;
        n5 = c.bxf[pi] * g5;
        n6 = c.byf[pj] * g6;
    }
    if (fw_in(c, 4, i, j)) {
        const T rb = a.rb[ij];
        const T dx = (T(27) * sxy[ij + P] - T(27) * sxy[ij] - sxy[ij + 2 * P] + sxy[ij - P]) * c.ihx;
        const T dy = (T(27) * syy[ij + 1] - T(27) * syy[ij] - syy[ij + 2] + syy[ij - 1]) * c.ihy;
        const T S4 = dx + msn(M7) + dy + msn(M8);
        const T g = a.A.f(VY, s)[ij];
        gB += -g * S4 * c.dt / (rb * rb);
        const T e = g * c.dt / rb;
        const T g7 = a7[ij] + e, g8 = a8[ij] + e;
        t2 = e + c.axh[pi] * g7;
        t3 = e + c.ayh[pj] * g8;
        n7 = c.bxh[pi] * g7;
        n8 = c.byh[pj] * g8;
    }
    a5[ij] = n5;
    a6[ij] = n6;
    a7[ij] = n7;
    a8[ij] = n8;
    T *tb = a.Tb + (int64_t)s * a.npad + ij;
    tb[0] = t0;
    tb[a.tstride] = t1;
    tb[2 * a.tstride] = t2;
    tb[3 * a.tstride] = t3;
}

// velocity phase gathers at (i, j): sxx <- (D-x)^T t0 ; sxy <- (D-y)^T t1 + (D+x)^T t2 ;
// syy <- (D+y)^T t3
template <typename T>
__device__ inline void adjv_b_cell(const ElAdj<T> &a, const ElCoef<T> &c, int s, int i, int j,
                                   bool from_tb = false)
{
    const int64_t ij = (int64_t)i * c.P + j;
    auto t = [&](int k, int ii, int jj) {
        if (!from_tb) return tv_vel(a, c, k, s, ii, jj);
        return ii >= 1 && ii <= c.NX && jj >= 1 && jj <= c.NY ? tb_at(a, k, s, (int64_t)ii * c.P + jj) : T(0);
    };
    const T gxx = (T(27) * t(0, i, j) - T(27) * t(0, i + 1, j) - t(0, i - 1, j) + t(0, i + 2, j)) * c.ihx;
    const T gxy = (T(27) * t(1, i, j) - T(27) * t(1, i, j + 1) - t(1, i, j - 1) + t(1, i, j + 2)) * c.ihy +
                  (T(27) * t(2, i - 1, j) - T(27) * t(2, i, j) - t(2, i - 2, j) + t(2, i + 1, j)) * c.ihx;
    const T gyy = (T(27) * t(3, i, j - 1) - T(27) * t(3, i, j) - t(3, i, j - 2) + t(3, i, j + 1)) * c.ihy;
    a.A.f(SXX, s)[ij] += gxx;
    a.A.f(SXY, s)[ij] += gxy;
    a.A.f(SYY, s)[ij] += gyy;
}

// stress phase, frame_a: fw1/fw2 transposed through m1..m4
template <typename T>
__device__ inline void adjs_a_cell(const ElAdj<T> &a, const ElCoef<T> &c, int s, int i, int j,
                                   T &gL, T &gM1, T &gM2)
{
    const int P = c.P;
    const int64_t ij = (int64_t)i * P + j;
    const T *__restrict__ vx = a.Sm.f(VX, s);
    const T *__restrict__ vy = a.Sm.f(VY, s);
    T *a1 = a.A.f(M1, s), *a2 = a.A.f(M2, s), *a3 = a.A.f(M3, s), *a4 = a.A.f(M4, s);
    T t0 = 0, t1 = 0, t2 = 0, t3 = 0, n1 = 0, n2 = 0, n3 = 0, n4 = 0;
    const int64_t mi = a.Sn.mv(i, j, ij);  // forward memory variables of level n (< 0: all 0)
    auto msn = [&](int q) { return mi >= 0 ? a.Sn.f(q, s)[mi] : T(0); };
    const int pi = i - 1 - c.sub, pj = j - 1 - c.sub;  // profile index
    if (fw_in(c, 1, i, j)) {
        const T lb = c.sub ? a.lam[ij] : T(0.5) * (a.lam[ij + P] + a.lam[ij]);
        const T mb = c.sub ? a.mu[ij] : T(0.5) * (a.mu[ij + P] + a.mu[ij]);
        const T lm = lb + T(2) * mb;
        const T dx = (T(27) * vx[ij + P] - T(27) * vx[ij] - vx[ij + 2 * P] + vx[ij - P]) * c.ihx;
        const T dy = (T(27) * vy[ij] - T(27) * vy[ij - 1] - vy[ij + 1] + vy[ij - 2]) * c.ihy;
        const T X = dx + msn(M1), Y = dy + msn(M2);
        const T gx = a.A.f(SXX, s)[ij] * c.dt, gy = a.A.f(SYY, s)[ij] * c.dt;
        const T glm = X * gx + Y * gy;
        gL += Y * gx + X * gy + glm;  // d/d lambda-bar
        gM1 += T(2) * glm;            // d/d mu-bar (fw1)
        const T gX = lm * gx + lb * gy, gY = lb * gx + lm * gy;
        const T g1 = a1[ij] + gX, g2 = a2[ij] + gY;
        t0 = gX + c.axh[pi] * g1;
        t1 = gY + c.ayf[pj] * g2;
        n1 = c.bxh[pi] * g1;
        n2 = c.byf[pj] * g2;
    }
    if (fw_in(c, 2, i, j)) {
        const T mb = c.sub ? a.mu[ij] : T(0.5) * (a.mu[ij] + a.mu[ij + 1]);
        const T dx = (T(27) * vy[ij] - T(27) * vy[ij - P] - vy[ij + P] + vy[ij - 2 * P]) * c.ihx;
        const T dy = (T(27) * vx[ij + 1] - T(27) * vx[ij] - vx[ij + 2] + vx[ij - 1]) * c.ihy;
        const T g = a.A.f(SXY, s)[ij];
        gM2 += g * (dx + msn(M3) + dy + msn(M4)) * c.dt;  // d/d mu-bar (fw2)
        const T gs = g * mb * c.dt;
        const T g3 = a3[ij] + gs, g4 = a4[ij] + gs;
        t2 = gs + c.axf[pi] * g3;
        t3 = gs + c.ayh[pj] * g4;
        n3 = c.bxf[pi] * g3;
        n4 = c.byh[pj] * g4;
    }
    a1[ij] = n1;
    a2[ij] = n2;
    a3[ij] = n3;
    a4[ij] = n4;
    T *tb = a.Tb + (int64_t)s * a.npad + ij;
    tb[0] = t0;
    tb[a.tstride] = t1;
    tb[2 * a.tstride] = t2;
    tb[3 * a.tstride] = t3;
}

// stress phase gathers: vx <- (D+x)^T t0 + (D+y)^T t3 ; vy <- (D-y)^T t1 + (D-x)^T t2
template <typename T>
__device__ inline void adjs_b_cell(const ElAdj<T> &a, const ElCoef<T> &c, int s, int i, int j,
                                   bool from_tb = false)
{
    const int64_t ij = (int64_t)i * c.P + j;
    auto t = [&](int k, int ii, int jj) {
        if (!from_tb) return tv_str(a, c, k, s, ii, jj);
        return ii >= 1 && ii <= c.NX && jj >= 1 && jj <= c.NY ? tb_at(a, k, s, (int64_t)ii * c.P + jj) : T(0);
    };
    const T gvx = (T(27) * t(0, i - 1, j) - T(27) * t(0, i, j) - t(0, i - 2, j) + t(0, i + 1, j)) * c.ihx +
                  (T(27) * t(3, i, j - 1) - T(27) * t(3, i, j) - t(3, i, j - 2) + t(3, i, j + 1)) * c.ihy;
    const T gvy = (T(27) * t(1, i, j) - T(27) * t(1, i, j + 1) - t(1, i, j - 1) + t(1, i, j + 2)) * c.ihy +
                  (T(27) * t(2, i, j) - T(27) * t(2, i + 1, j) - t(2, i - 1, j) + t(2, i + 2, j)) * c.ihx;
    a.A.f(VX, s)[ij] += gvx;
    a.A.f(VY, s)[ij] += gvy;
}

// frame kernels: one column per lane, FRPT rows per wavefront item, grid y = shot group
template <typename T>
__device__ inline void adjv_frame_a_body(const ElAdj<T> &a, const ElCoef<T> &c,
                                         const int4 *__restrict__ items, int nitems, int bx, int g)
{
    const int it = bx * NWAVES + wave_id();
    if (it >= nitems) return;
    const int4 m = items[it];
    const int j = m.z + (threadIdx.x & (WAVE - 1));
    if (j >= m.w) return;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int r = 0; r < m.y; r++) {
        const int i = m.x + r;
        T gR = 0, gB = 0;
        for (int s = g * a.Sg; s < s1; s++) adjv_a_cell(a, c, s, i, j, gR, gB);
        const int64_t ij = (int64_t)i * c.P + j;
        g_at(a, G_RHO, g, ij) += gR;
        g_at(a, G_RB, g, ij) += gB;
    }
}
// Stress phase, ring of the rectangle (its cells within two of its edge, which the frame
// cells' gathers reach): t0..t3 recomputed as on rectangle cells and stored in the
// temporaries, so that frame_b reads every neighbour's t-value from the temporaries alone.
// One column per lane.
template <typename T>
__device__ inline void adjs_ring_body(const ElAdj<T> &a, const ElCoef<T> &c,
                                      const int4 *__restrict__ items, int nitems, int bx, int g)
{
    const int it = bx * NWAVES + wave_id();
    if (it >= nitems) return;
    const int4 m = items[it];
    const int j = m.z + (threadIdx.x & (WAVE - 1));
    if (j >= m.w) return;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int r = 0; r < m.y; r++) {
        const int i = m.x + r;
        const int64_t ij = (int64_t)i * c.P + j;
        for (int s = g * a.Sg; s < s1; s++) {
            T *tb = a.Tb + (int64_t)s * a.npad + ij;
            tb[0] = tv_str(a, c, 0, s, i, j);
            tb[a.tstride] = tv_str(a, c, 1, s, i, j);
            const T t2 = tv_str(a, c, 2, s, i, j);
            tb[2 * a.tstride] = t2;
            tb[3 * a.tstride] = t2;
        }
    }
}
// frame_a (and, in the stress phase, the ring) in one launch: blocks [0, nfa) frame items,
// then ring items (nib and nrb item blocks per shot group)
template <typename T>
__global__ __launch_bounds__(WAVE *NWAVES) void el_adjv_a(ElAdj<T> a, ElCoef<T> c, const int4 *__restrict__ items,
                                                          int nitems, int nib, const int4 *__restrict__,
                                                          int, int, int nfa)
{
    const int b = (int)blockIdx.x;
    if (b < nfa) adjv_frame_a_body(a, c, items, nitems, b % nib, b / nib);
}
template <typename T>
__global__ __launch_bounds__(WAVE *NWAVES) void el_adjs_a(ElAdj<T> a, ElCoef<T> c, const int4 *__restrict__ items,
                                                          int nitems, int nib, const int4 *__restrict__ ring,
                                                          int nring, int nrb, int nfa)
{
    const int b = (int)blockIdx.x;
    if (b < nfa) adjs_frame_a_body(a, c, items, nitems, b % nib, b / nib);
    else adjs_ring_body(a, c, ring, nring, (b - nfa) % nrb, (b - nfa) / nrb);
}

template <typename T>
__device__ inline void adjv_frame_b_body(const ElAdj<T> &a, const ElCoef<T> &c,
                                         const int4 *__restrict__ items, int nitems, int bx, int g)
{
    const int it = bx * NWAVES + wave_id();
    if (it >= nitems) return;
    const int4 m = items[it];
    const int j = m.z + (threadIdx.x & (WAVE - 1));
    if (j >= m.w) return;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int r = 0; r < m.y; r++)
        for (int s = g * a.Sg; s < s1; s++) adjv_b_cell(a, c, s, m.x + r, j);
}
template <typename T>
__global__ __launch_bounds__(WAVE *NWAVES) void el_adjv_frame_b(ElAdj<T> a, ElCoef<T> c,
                                                                const int4 *__restrict__ items, int nitems)
{
    adjv_frame_b_body(a, c, items, nitems, (int)blockIdx.x, (int)blockIdx.y);
}
template <typename T>
__device__ inline void adjs_frame_a_body(const ElAdj<T> &a, const ElCoef<T> &c,
                                         const int4 *__restrict__ items, int nitems, int bx, int g)
{
    const int it = bx * NWAVES + wave_id();
    if (it >= nitems) return;
    const int4 m = items[it];
    const int j = m.z + (threadIdx.x & (WAVE - 1));
    if (j >= m.w) return;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int r = 0; r < m.y; r++) {
        const int i = m.x + r;
        T gL = 0, gM1 = 0, gM2 = 0;
        for (int s = g * a.Sg; s < s1; s++) adjs_a_cell(a, c, s, i, j, gL, gM1, gM2);
        const int64_t ij = (int64_t)i * c.P + j;
        g_at(a, G_LB, g, ij) += gL;
        g_at(a, G_MB1, g, ij) += gM1;
        g_at(a, G_MB2, g, ij) += gM2;
    }
}
template <typename T>
__device__ inline void adjs_frame_b_body(const ElAdj<T> &a, const ElCoef<T> &c,
                                         const int4 *__restrict__ items, int nitems, int bx, int g)
{
    const int it = bx * NWAVES + wave_id();
    if (it >= nitems) return;
    const int4 m = items[it];
    const int j = m.z + (threadIdx.x & (WAVE - 1));
    if (j >= m.w) return;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int r = 0; r < m.y; r++)
        for (int s = g * a.Sg; s < s1; s++) adjs_b_cell(a, c, s, m.x + r, j, true);
}
template <typename T>
__global__ __launch_bounds__(WAVE *NWAVES) void el_adjs_frame_b(ElAdj<T> a, ElCoef<T> c,
                                                                const int4 *__restrict__ items, int nitems)
{
    adjs_frame_b_body(a, c, items, nitems, (int)blockIdx.x, (int)blockIdx.y);
}

// Rectangle kernels: a workgroup owns an LT_I x LT_J tile and loops over its shot group.  Per
// shot it stages in LDS, for the tile and a 2-cell halo, t0..t3 (recomputed from the field
// cotangents on rectangle cells, read from the temporaries on frame cells) and, in the velocity
// phase, the three stresses the parameter terms differentiate; then each thread takes four
// cells of one column.  (Staging vx, vy for the stress phase measured slower: 137 -> 156 us
// per level at 600 x 600 x 16 shots, the larger LDS tile lowering occupancy.)
constexpr int LT_I = 16, LT_J = 64, LH = 2;
constexpr int LS_I = LT_I + 2 * LH, LS_J = LT_J + 2 * LH;

template <typename T> __device__ inline T d4(T a0, T a1, T am, T ap)  // 27 a0 - 27 a1 - am + ap
{
    return T(27) * a0 - T(27) * a1 - am + ap;
}

template <typename T>
__device__ inline void adjv_rect_body(const ElAdj<T> &a, const ElCoef<T> &c, const TileMap &map, int bid)
{
    __shared__ T S[7][LS_I][LS_J];  // t0..t3, then sxx, sxy, syy of level n
    int g, strip, tj;
    if (!decode_tile(map, g, strip, tj, bid)) return;
    const int i0 = a.R.iL + strip * LT_I, j0 = a.R.jL + tj * LT_J;
    const int tc = threadIdx.x % LT_J, tr = threadIdx.x / LT_J;
    const int j = j0 + tc;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    T gR[4] = {}, gB[4] = {}, irho[4], irb[4];
    bool own[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {  // shot-invariant: 1/rho, 1/rho-bar of this thread's cells
        const int i = i0 + tr + 4 * r;
        own[r] = i < a.R.iR && j < a.R.jR;
        const int64_t ij = (int64_t)i * c.P + j;
        irho[r] = own[r] ? T(1) / a.rho[ij] : T(0);
        irb[r] = own[r] ? T(1) / a.rb[ij] : T(0);
    }
    for (int s = g * a.Sg; s < s1; s++) {
        // this shot's stress cotangents at the thread's cells, loaded before the staging so
        // that their latency overlaps it
        T axx[4], axy[4], ayy[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t ij = (int64_t)(i0 + tr + 4 * r) * c.P + j;
            axx[r] = own[r] ? a.A.f(SXX, s)[ij] : T(0);
            axy[r] = own[r] ? a.A.f(SXY, s)[ij] : T(0);
            ayy[r] = own[r] ? a.A.f(SYY, s)[ij] : T(0);
        }
        for (int q = threadIdx.x; q < LS_I * LS_J; q += 256) {
            const int li = q / LS_J, lj = q % LS_J;
            const int i = i0 - LH + li, jj = j0 - LH + lj;
            T t[4] = {T(0), T(0), T(0), T(0)}, f[3] = {T(0), T(0), T(0)};
            if (i >= 0 && i <= c.NX + 1 && jj >= 0 && jj <= c.NY + 1) {
                const int64_t ij = (int64_t)i * c.P + jj;
                f[0] = a.Sn.f(SXX, s)[ij];
                f[1] = a.Sn.f(SXY, s)[ij];
                f[2] = a.Sn.f(SYY, s)[ij];
                if (i >= 1 && i <= c.NX && jj >= 1 && jj <= c.NY) {
                    if (in_rect(a, i, jj)) {
                        t[0] = t[1] = a.A.f(VX, s)[ij] * c.dt / a.rho[ij];
                        t[2] = t[3] = a.A.f(VY, s)[ij] * c.dt / a.rb[ij];
                    } else {
#pragma unroll
                        for (int k = 0; k < 4; k++) t[k] = tb_at(a, k, s, ij);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 4; k++) S[k][li][lj] = t[k];
#pragma unroll
            for (int k = 0; k < 3; k++) S[4 + k][li][lj] = f[k];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int li = tr + 4 * r, i = i0 + li;
            if (!own[r]) continue;
            const int I = li + LH, J = tc + LH;
            const int64_t ij = (int64_t)i * c.P + j;
            // parameter terms (fw3: rho; fw4: rho-bar); t0 = g_vx dt / rho, t2 = g_vy dt / rho-bar
            const T S3 = d4(S[4][I][J], S[4][I - 1][J], S[4][I + 1][J], S[4][I - 2][J]) * c.ihx +
                         d4(S[5][I][J], S[5][I][J - 1], S[5][I][J + 1], S[5][I][J - 2]) * c.ihy;
            const T S4 = d4(S[5][I + 1][J], S[5][I][J], S[5][I + 2][J], S[5][I - 1][J]) * c.ihx +
                         d4(S[6][I][J + 1], S[6][I][J], S[6][I][J + 2], S[6][I][J - 1]) * c.ihy;
            gR[r] -= S[0][I][J] * S3 * irho[r];
            gB[r] -= S[2][I][J] * S4 * irb[r];
            // sxx <- (D-x)^T t0 ; sxy <- (D-y)^T t1 + (D+x)^T t2 ; syy <- (D+y)^T t3
            const T gxx = d4(S[0][I][J], S[0][I + 1][J], S[0][I - 1][J], S[0][I + 2][J]) * c.ihx;
            const T gxy = d4(S[1][I][J], S[1][I][J + 1], S[1][I][J - 1], S[1][I][J + 2]) * c.ihy +
                          d4(S[2][I - 1][J], S[2][I][J], S[2][I - 2][J], S[2][I + 1][J]) * c.ihx;
            const T gyy = d4(S[3][I][J - 1], S[3][I][J], S[3][I][J - 2], S[3][I][J + 1]) * c.ihy;
            a.A.f(SXX, s)[ij] = axx[r] + gxx;
            a.A.f(SXY, s)[ij] = axy[r] + gxy;
            a.A.f(SYY, s)[ij] = ayy[r] + gyy;
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = i0 + tr + 4 * r;
        if (i >= a.R.iR || j >= a.R.jR) continue;
        const int64_t ij = (int64_t)i * c.P + j;
        g_at(a, G_RHO, g, ij) += gR[r];
        g_at(a, G_RB, g, ij) += gB[r];
    }
}

template <typename T>
__device__ inline void adjs_rect_body(const ElAdj<T> &a, const ElCoef<T> &c, const TileMap &map, int bid)
{
    __shared__ T S[4][LS_I][LS_J];  // t0..t3
    int g, strip, tj;
    if (!decode_tile(map, g, strip, tj, bid)) return;
    const int i0 = a.R.iL + strip * LT_I, j0 = a.R.jL + tj * LT_J;
    const int tc = threadIdx.x % LT_J, tr = threadIdx.x / LT_J;
    const int j = j0 + tc;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    T gL[4] = {}, gM1[4] = {}, gM2[4] = {};
    bool own[4];
#pragma unroll
    for (int r = 0; r < 4; r++) own[r] = i0 + tr + 4 * r < a.R.iR && j < a.R.jR;
    for (int s = g * a.Sg; s < s1; s++) {
        // this shot's cotangents at the thread's cells, loaded before the staging so that
        // their latency overlaps it
        T avx[4], avy[4], axx[4], ayy[4], axy[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t ij = (int64_t)(i0 + tr + 4 * r) * c.P + j;
            avx[r] = own[r] ? a.A.f(VX, s)[ij] : T(0);
            avy[r] = own[r] ? a.A.f(VY, s)[ij] : T(0);
            axx[r] = own[r] ? a.A.f(SXX, s)[ij] : T(0);
            ayy[r] = own[r] ? a.A.f(SYY, s)[ij] : T(0);
            axy[r] = own[r] ? a.A.f(SXY, s)[ij] : T(0);
        }
        for (int q = threadIdx.x; q < LS_I * LS_J; q += 256) {
            const int li = q / LS_J, lj = q % LS_J;
            const int i = i0 - LH + li, jj = j0 - LH + lj;
            T t[4] = {T(0), T(0), T(0), T(0)};
            if (i >= 1 && i <= c.NX && jj >= 1 && jj <= c.NY) {
                const int64_t ij = (int64_t)i * c.P + jj;
                if (in_rect(a, i, jj)) {
#pragma unroll
                    for (int k = 0; k < 3; k++) t[k] = tv_str(a, c, k, s, i, jj);
                    t[3] = t[2];
                } else {
#pragma unroll
                    for (int k = 0; k < 4; k++) t[k] = tb_at(a, k, s, ij);
                }
            }
#pragma unroll
            for (int k = 0; k < 4; k++) S[k][li][lj] = t[k];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int li = tr + 4 * r, i = i0 + li;
            if (!own[r]) continue;
            const int I = li + LH, J = tc + LH, P = c.P;
            const int64_t ij = (int64_t)i * P + j;
            // parameter terms: fw1 (lambda-bar, mu-bar) and fw2 (mu-bar), from level n-1
            const T *__restrict__ vx = a.Sm.f(VX, s);
            const T *__restrict__ vy = a.Sm.f(VY, s);
            const T X = d4(vx[ij + P], vx[ij], vx[ij + 2 * P], vx[ij - P]) * c.ihx;
            const T Y = d4(vy[ij], vy[ij - 1], vy[ij + 1], vy[ij - 2]) * c.ihy;
            const T gx = axx[r] * c.dt, gy = ayy[r] * c.dt;
            const T glm = X * gx + Y * gy;
            gL[r] += Y * gx + X * gy + glm;
            gM1[r] += T(2) * glm;
            const T dx = d4(vy[ij], vy[ij - P], vy[ij + P], vy[ij - 2 * P]) * c.ihx;
            const T dy = d4(vx[ij + 1], vx[ij], vx[ij + 2], vx[ij - 1]) * c.ihy;
            gM2[r] += axy[r] * (dx + dy) * c.dt;
            // vx <- (D+x)^T t0 + (D+y)^T t3 ; vy <- (D-y)^T t1 + (D-x)^T t2
            const T gvx = d4(S[0][I - 1][J], S[0][I][J], S[0][I - 2][J], S[0][I + 1][J]) * c.ihx +
                          d4(S[3][I][J - 1], S[3][I][J], S[3][I][J - 2], S[3][I][J + 1]) * c.ihy;
            const T gvy = d4(S[1][I][J], S[1][I][J + 1], S[1][I][J - 1], S[1][I][J + 2]) * c.ihy +
                          d4(S[2][I][J], S[2][I + 1][J], S[2][I - 1][J], S[2][I + 2][J]) * c.ihx;
            a.A.f(VX, s)[ij] = avx[r] + gvx;
            a.A.f(VY, s)[ij] = avy[r] + gvy;
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = i0 + tr + 4 * r;
        if (i >= a.R.iR || j >= a.R.jR) continue;
        const int64_t ij = (int64_t)i * c.P + j;
        g_at(a, G_LB, g, ij) += gL[r];
        g_at(a, G_MB1, g, ij) += gM1[r];
        g_at(a, G_MB2, g, ij) += gM2[r];
    }
}

// Rectangle kernels alone, and fused with frame_b (independent cells, same phase): workgroups
// [0, nfb) take the frame items (nib item blocks per shot group), the rest the rectangle tiles.
template <typename T>
__global__ __launch_bounds__(256) void el_adjv_rect(ElAdj<T> a, ElCoef<T> c, TileMap map)
{
    adjv_rect_body(a, c, map, (int)blockIdx.x);
}
template <typename T>
__global__ __launch_bounds__(256) void el_adjs_rect(ElAdj<T> a, ElCoef<T> c, TileMap map)
{
    adjs_rect_body(a, c, map, (int)blockIdx.x);
}
template <typename T>
__global__ __launch_bounds__(256) void el_adjv_rb(ElAdj<T> a, ElCoef<T> c, TileMap map,
                                                  const int4 *__restrict__ items, int nitems, int nib,
                                                  int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) adjv_frame_b_body(a, c, items, nitems, b % nib, b / nib);
    else adjv_rect_body(a, c, map, b - nfb);
}
template <typename T>
__global__ __launch_bounds__(256) void el_adjs_rb(ElAdj<T> a, ElCoef<T> c, TileMap map,
                                                  const int4 *__restrict__ items, int nitems, int nib,
                                                  int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) adjs_frame_b_body(a, c, items, nitems, b % nib, b / nib);
    else adjs_rect_body(a, c, map, b - nfb);
}

// ---- streaming rectangle adjoint ----
// A workgroup is SW_NW wavefronts, one shot each (one shot group), over one strip of SW_COLS
// output columns and one chunk of up to `ch` output rows of the rectangle.  Lane l holds column
// jL + strip*SW_COLS - 2 + l; lanes 2 .. WAVE-3 are the output columns, the two lanes on each
// side carry the halo columns of the y-direction stencils.  Each wavefront walks its chunk row
// by row, keeping the x-direction stencil inputs of rows i-2 .. i+2 in registers (one new row
// loaded per step) and taking the y-direction neighbours of row i from the adjacent lanes.
// t-values inside the rectangle are formed from the cotangents (t0 = t1, t2 = t3 there); on
// the frame cells the halo reaches they come from the temporaries frame_a wrote, and outside
// the interior they are 0, as in tv_vel / tv_str.  Each wavefront writes its shot's parameter
// terms per cell to its own LDS slice; after one barrier the slices are added in wavefront
// order (no atomics: the sums, like every other sum of the plan, are bitwise reproducible) and
// added to the group's accumulators once per workgroup.  LDS: shots x terms x rows x WAVE.
constexpr int SW_NW = 8, SW_CH_MAX = 32, SW_COLS = WAVE - 4, SW_LDS_ROWS = 64;  // shots x rows <= 64

struct StrmMap {
    int nstrip, nchunk, ngroup, nb, ch;
};
__device__ inline bool decode_strm(const StrmMap &m, int &g, int &chunk, int &strip, int b)
{
    const int nb8 = (m.nb + NXCD - 1) / NXCD;
    const int L = (b % NXCD) * nb8 + b / NXCD;
    if (L >= m.nb) return false;
    strip = L % m.nstrip;
    const int r = L / m.nstrip;
    chunk = r % m.nchunk;
    g = r / m.nchunk;
    return true;
}
template <typename T> __device__ inline T lane_at(T v, int src) { return __shfl(v, src, WAVE); }
// Loads and stores at a 32-bit byte offset from a wave-uniform base: the base stays in scalar
// registers and every load of one row shares the row's offset register (one field of one
// shot is below 4 GiB; the host checks it before choosing these kernels).
template <typename T> __device__ inline T ld_at(const T *base, uint32_t off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T> __device__ inline void st_at(T *base, uint32_t off, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}

// shot-invariant inputs of the streaming kernels (workspace arrays, one Npad grid each)
template <typename T> struct ElPar {
    const T *irho, *irb, *lb, *mb, *mb2;  // 1/rho, 1/rho-bar, lambda-bar, mu-bar (fw1), mu-bar (fw2)
};

// Each row step issues the loads of the row it will need one step later (window row i+3 and
// the row-(i+1) cotangents), so their latency overlaps the current row's arithmetic.

// Full-grid streaming form: the same walk over every cell of the padded grid, the CPML frame
// included, so no frame_a / frame_b pass and no temporaries.  t-values are formed in the
// registers of the lanes that need them: t0 and t2 when a row enters the window, t1 and t3 on
// the output row, each with the memory-variable cotangent term (a * (m + e), the profile a 0
// outside the CPML strips) and 0 outside the cell's update range (as tv_vel / tv_str).  The
// memory-variable cotangents are read from the A arrays and their updates written to the
// second set (mo), so a halo lane never reads a value another wavefront has already replaced;
// the sets swap every level.  A memory-variable cotangent whose a-profile is 0 at its row (x)
// or column (y) never reaches a t-value, so it is neither read nor written there.  A
// wavefront whose window rows and lanes all lie inside R (identity CPML, every update range)
// runs the plain specialisation of the walk: no range tests, no memory variables.
template <typename T> struct ProfRow {  // CPML coefficients of a row (x) or column (y), 0 off range
    T af, ah, bf, bh;
};
template <typename T> __device__ inline ProfRow<T> prof_row(const ElCoef<T> &c, int r)  // r in [-2, NX+3]
{
    const T *p = c.prow + 4 * (r + 2);
    return ProfRow<T>{p[0], p[1], p[2], p[3]};
}
template <typename T> __device__ inline ProfRow<T> prof_col(const ElCoef<T> &c, int j)  // j in [0, NY+1]
{
    const T *p = c.pcol + 4 * j;
    return ProfRow<T>{p[0], p[1], p[2], p[3]};
}
// The update ranges of fw_in split into row and column parts: fw1 = A_row & A_col, fw2 = B_row
// & B_col, fw3 = B_row & A_col, fw4 = A_row & B_col.
struct FwPart {
    bool A, B;
};
template <int SUB, typename T> __device__ inline FwPart fw_row(const ElCoef<T> &c, int i)
{
    if constexpr (SUB) return FwPart{i >= 2 && i <= c.NX - 1, i >= 2 && i <= c.NX - 1};
    return FwPart{i >= 1 && i <= c.NX - 1, i >= 2 && i <= c.NX};
}
template <int SUB, typename T> __device__ inline FwPart fw_col(const ElCoef<T> &c, int j)
{
    if constexpr (SUB) return FwPart{j >= 2 && j <= c.NY - 1, j >= 2 && j <= c.NY - 1};
    return FwPart{j >= 2 && j <= c.NY, j >= 1 && j <= c.NY - 1};
}
// Index of (i, j) among the cells outside the box mb (FrameBox::index) as a row part and a
// column part: rows above and below the box are whole rows, the rows beside it keep the qj0
// columns to its left and the W - qj1 to its right.  32-bit: one field of one shot.
struct MvCol {
    int full, mid;  // the column's offset in a whole row / in a row beside the box (-1: inside)
};
__device__ inline MvCol mv_col(const FrameBox &b, int j)
{
    return MvCol{j, j < b.qj0 ? j : (j >= b.qj1 ? b.qj0 + (j - b.qj1) : -1)};
}
__device__ inline int mv_at(const FrameBox &b, const MvCol &cj, int i)
{
    const int top = b.qi0 * b.W, wm = b.qj0 + (b.W - b.qj1);
    if (i < b.qi0) return i * b.W + cj.full;
    if (i >= b.qi1) return top + (b.qi1 - b.qi0) * wm + (i - b.qi1) * b.W + cj.full;
    return cj.mid < 0 ? -1 : top + (i - b.qi0) * wm + cj.mid;
}

// ---- fused streaming forward step ----
// One launch per level over the whole padded grid: fw1/fw2 and then fw3/fw4 of the same rows,
// the new stresses kept in registers for the velocity update instead of a round trip through
// HBM.  A workgroup is NWAVES wavefronts, one shot each, over one strip of FS_COLS output
// columns and one chunk of up to ch output rows.  Lane l holds column jb + l, jb = strip *
// FS_COLS - 4: the stresses of level n are formed on lanes 2 .. WAVE-3 (their y-direction
// velocity neighbours come from the adjacent lanes), the velocities on the output lanes
// 4 .. WAVE-5 (their stress neighbours from lanes 2 .. WAVE-3).  The wavefront walks rows
// r = i0-2 .. i1+1: the stresses of row r from the level n-1 velocities of rows r-2 .. r+2 in
// a register window, then the velocities of row r-2 from the new stresses of rows r-4 .. r in
// a second window.  The stresses of the two rows above and below the chunk and of the halo
// lanes are formed again, with the same arithmetic, by the wavefronts that own them, and only
// the owner stores them.  A wavefront whose rows and lanes all lie inside R runs the plain
// walk (no range tests, no memory variables; the CPML is the identity there).  Sources and
// receivers stay in el_src_rcv after the step: the velocity update reads the stresses as
// they are before the stress sources are added.
constexpr int FS_COLS = WAVE - 8;

// Lane shifts as DPP wavefront shifts (VALU moves, no LDS): up1 gives lane l the value of lane
// l + 1 (wave_shl:1), dn1 that of lane l - 1 (wave_shr:1); a lane shifted in from beyond the
// wavefront reads 0.  DPP reads an inactive source lane as that 0 too, so the walks call them
// where every lane of the wavefront is active (the builtin is convergent: the compiler keeps
// it there rather than sinking it into the per-lane branches that use the result).
template <int CTRL> __device__ inline float dpp_move(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL> __device__ inline double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <typename T> __device__ inline T up1(T v) { return dpp_move<0x130>(v); }
template <typename T> __device__ inline T dn1(T v) { return dpp_move<0x138>(v); }

template <int SUB, typename T>
__device__ inline void fwd_full_body(const ElC<T> &I, const El<T> &O, const T *__restrict__ rho,
                                     const T *__restrict__ lam, const T *__restrict__ mu,
                                     const ElCoef<T> &c, const Rect &R, const StrmMap &mp, int act,
                                     int bid)
{
    int g, chunk, strip;
    if (!decode_strm(mp, g, chunk, strip, bid)) return;
    const int s = g * NWAVES + wave_id();
    if (s >= act) return;  // wave-uniform; the kernel has no barrier
    const int lane = threadIdx.x % WAVE;
    const int NXp = c.NX + 1, NYp = c.NY + 1;  // last padded row / column
    const int i0 = chunk * mp.ch, i1 = min(i0 + mp.ch, NXp + 1);
    const int jb = strip * FS_COLS - 4, j = jb + lane;
    const int jc = j < 0 ? 0 : (j > NYp ? NYp : j);  // loads of lanes beyond the grid stay inside it
    const bool own = lane >= 4 && lane < WAVE - 4 && j <= NYp;
    const bool sl = lane >= 2 && lane < WAVE - 2;  // lanes whose stresses are used
    const bool gen = !(i0 - 4 >= 0 && i1 + 3 <= NXp && i0 - 2 >= R.iL && i1 + 2 <= R.iR &&
                       jb + 2 >= R.jL && jb + WAVE - 2 <= R.jR);
    const T dt = c.dt, ihx = c.ihx, ihy = c.ihy;
    const uint32_t pb = (uint32_t)c.P * (uint32_t)sizeof(T), jb4 = (uint32_t)jc * (uint32_t)sizeof(T);
    const T *__restrict__ vx = I.f(VX, s);
    const T *__restrict__ vy = I.f(VY, s);
    const T *__restrict__ ixx = I.f(SXX, s);
    const T *__restrict__ iyy = I.f(SYY, s);
    const T *__restrict__ ixy = I.f(SXY, s);
    T *__restrict__ ovx = O.f(VX, s);
    T *__restrict__ ovy = O.f(VY, s);
    T *__restrict__ oxx = O.f(SXX, s);
    T *__restrict__ oyy = O.f(SYY, s);
    T *__restrict__ oxy = O.f(SXY, s);
    auto walk = [&](auto genc) {
        constexpr bool GEN = decltype(genc)::value;
        const T *im[8];  // memory variables of level n-1 and n (general walk only)
        T *om[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            im[q] = GEN ? I.f(M1 + q, s) : nullptr;
            om[q] = GEN ? O.f(M1 + q, s) : nullptr;
        }
        const ProfRow<T> py = GEN ? prof_col(c, jc) : ProfRow<T>{};
        const FwPart fc = GEN ? fw_col<SUB>(c, j) : FwPart{true, true};
        const MvCol mci = mv_col(I.mb, jc), mco = mv_col(O.mb, jc);
        // memory-variable index of (r, jc): compact outside the box, or the full grid
        auto mvi = [&](const FrameBox &b, const MvCol &m, int r) -> int {
            return b.H ? mv_at(b, m, r) : r * c.P + jc;
        };
        // row r of an array (0 off the grid's rows in the general walk)
        auto row = [&](const T *__restrict__ f, int r) -> T {
            const bool ok = !GEN || (r >= 0 && r <= NXp);  // uniform over the wavefront
            return ok ? ld_at(f, (uint32_t)(ok ? r : 0) * pb + jb4) : T(0);
        };
        // the loads of step r: the window rows entering, the level n-1 stresses (and memory
        // variables m1..m4) of row r, those of m5..m8 of row r-2
        struct Rw { T x, y, l, u, xx, yy, xy, rh, m1, m2, m3, m4, m5, m6, m7, m8; };
        auto fetch = [&](int r, Rw &o) {
            o.x = row(vx, r + 2);
            o.y = row(vy, r + 1);
            o.l = row(lam, r + 1);
            o.u = row(mu, r + 1);
            o.xx = row(ixx, r);
            o.yy = row(iyy, r);
            o.xy = row(ixy, r);
            o.rh = row(rho, r - 1);
            o.m1 = o.m2 = o.m3 = o.m4 = o.m5 = o.m6 = o.m7 = o.m8 = T(0);
            if constexpr (GEN) {
                const int mi = sl && r >= 0 && r <= NXp ? mvi(I.mb, mci, r) : -1;
                if (mi >= 0) {
                    const uint32_t o4 = (uint32_t)mi * (uint32_t)sizeof(T);
                    o.m1 = ld_at(im[0], o4);
                    o.m2 = ld_at(im[1], o4);
                    o.m3 = ld_at(im[2], o4);
                    o.m4 = ld_at(im[3], o4);
                }
                const int mv = own && r - 2 >= i0 ? mvi(I.mb, mci, r - 2) : -1;
                if (mv >= 0) {
                    const uint32_t o4 = (uint32_t)mv * (uint32_t)sizeof(T);
                    o.m5 = ld_at(im[4], o4);
                    o.m6 = ld_at(im[5], o4);
                    o.m7 = ld_at(im[6], o4);
                    o.m8 = ld_at(im[7], o4);
                }
            }
        };
        const int r0 = i0 - 2;
        // windows: WX rows r-2 .. r+2, WY rows r-2 .. r+1, LA / MU rows r, r+1, RH rows r-2,
        // r-1 (level n-1 and parameters); XX rows r-4 .. r, XY rows r-3 .. r, YY rows r-2 .. r
        // (new stresses)
        T WX[5], WY[4], LA[2], MU[2], RH[2];
        T XX[5] = {}, XY[4] = {}, YY[3] = {};
#pragma unroll
        for (int k = 0; k < 4; k++) WX[k] = row(vx, r0 - 2 + k);
#pragma unroll
        for (int k = 0; k < 3; k++) WY[k] = row(vy, r0 - 2 + k);
        LA[0] = row(lam, r0);
        MU[0] = row(mu, r0);
        RH[0] = row(rho, r0 - 2);
        Rw nx;
        fetch(r0, nx);
        for (int r = r0; r <= i1 + 1; r++) {
            const Rw cu = nx;
            if (r <= i1) fetch(r + 1, nx);
            WX[4] = cu.x;
            WY[3] = cu.y;
            LA[1] = cu.l;
            MU[1] = cu.u;
            RH[1] = cu.rh;
            // fw1 / fw2 on row r
            const T X = WX[2], Y = WY[2];
            const T vym = dn1(Y), vyp = up1(Y), vymm = dn1(vym);
            const T vxm = dn1(X), vxp = up1(X), vxpp = up1(vxp);
            const T mue = up1(MU[0]);
            const FwPart fr = GEN ? fw_row<SUB>(c, r) : FwPart{true, true};
            const ProfRow<T> px = GEN ? prof_row(c, r) : ProfRow<T>{};
            T sxx = cu.xx, syy = cu.yy, sxy = cu.xy;
            T m1 = T(0), m2 = T(0), m3 = T(0), m4 = T(0);
            if (!GEN || (fr.A && fc.A)) {
                const T lb = SUB ? LA[0] : T(0.5) * (LA[1] + LA[0]);
                const T mb = SUB ? MU[0] : T(0.5) * (MU[1] + MU[0]);
                const T lm = lb + T(2) * mb;
                T dx = (T(27) * WX[3] - T(27) * X - WX[4] + WX[1]) * ihx;
                T dy = (T(27) * Y - T(27) * vym - vyp + vymm) * ihy;
                if constexpr (GEN) {
                    m1 = px.bh * cu.m1 + px.ah * dx;
                    m2 = py.bf * cu.m2 + py.af * dy;
                    dx += m1;
                    dy += m2;
                }
                sxx += (lm * dx + lb * dy) * dt;
                syy += (lm * dy + lb * dx) * dt;
            }
            if (!GEN || (fr.B && fc.B)) {
                const T mb = SUB ? MU[0] : T(0.5) * (MU[0] + mue);
                const T dx = (T(27) * Y - T(27) * WY[1] - WY[3] + WY[0]) * ihx;
                const T dy = (T(27) * vxp - T(27) * X - vxpp + vxm) * ihy;
                if constexpr (GEN) {
                    m3 = px.bf * cu.m3 + px.af * dx;
                    m4 = py.bh * cu.m4 + py.ah * dy;
                    sxy += mb * (dx + m3 + dy + m4) * dt;
                } else {
                    sxy += mb * (dx + dy) * dt;
                }
            }
            if (own && r >= i0 && r < i1) {
                const uint32_t ij = (uint32_t)r * pb + jb4;
                st_at(oxx, ij, sxx);
                st_at(oyy, ij, syy);
                st_at(oxy, ij, sxy);
                if constexpr (GEN) {
                    const int mo = mvi(O.mb, mco, r);
                    if (mo >= 0) {
                        const uint32_t o4 = (uint32_t)mo * (uint32_t)sizeof(T);
                        st_at(om[0], o4, m1);
                        st_at(om[1], o4, m2);
                        st_at(om[2], o4, m3);
                        st_at(om[3], o4, m4);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 4; k++) XX[k] = XX[k + 1];
#pragma unroll
            for (int k = 0; k < 3; k++) XY[k] = XY[k + 1];
#pragma unroll
            for (int k = 0; k < 2; k++) YY[k] = YY[k + 1];
            XX[4] = sxx;
            XY[3] = sxy;
            YY[2] = syy;
            // fw3 / fw4 on row i = r - 2
            const int i = r - 2;
            if (i >= i0) {  // uniform
                const T B = XY[1], Q = YY[0];
                const T xym = dn1(B), xyp = up1(B), xymm = dn1(xym);
                const T yym = dn1(Q), yyp = up1(Q), yypp = up1(yyp);
                const T re0 = up1(RH[0]), re1 = up1(RH[1]);
                const FwPart fv = GEN ? fw_row<SUB>(c, i) : FwPart{true, true};
                const ProfRow<T> pv = GEN ? prof_row(c, i) : ProfRow<T>{};
                T vxn = WX[0], vyn = WY[0];
                T m5 = T(0), m6 = T(0), m7 = T(0), m8 = T(0);
                if (!GEN || (fv.B && fc.A)) {
                    const T dx = (T(27) * XX[2] - T(27) * XX[1] - XX[3] + XX[0]) * ihx;
                    const T dy = (T(27) * B - T(27) * xym - xyp + xymm) * ihy;
                    if constexpr (GEN) {
                        m5 = pv.bf * cu.m5 + pv.af * dx;
                        m6 = py.bf * cu.m6 + py.af * dy;
                        vxn += (dx + m5 + dy + m6) * dt / RH[0];
                    } else {
                        vxn += (dx + dy) * dt / RH[0];
                    }
                }
                if (!GEN || (fv.A && fc.B)) {
                    const T rb = SUB ? RH[0] : T(0.25) * (RH[0] + RH[1] + re1 + re0);
                    const T dx = (T(27) * XY[2] - T(27) * B - XY[3] + XY[0]) * ihx;
                    const T dy = (T(27) * yyp - T(27) * Q - yypp + yym) * ihy;
                    if constexpr (GEN) {
                        m7 = pv.bh * cu.m7 + pv.ah * dx;
                        m8 = py.bh * cu.m8 + py.ah * dy;
                        vyn += (dx + m7 + dy + m8) * dt / rb;
                    } else {
                        vyn += (dx + dy) * dt / rb;
                    }
                }
                if (own && i < i1) {
                    const uint32_t ij = (uint32_t)i * pb + jb4;
                    st_at(ovx, ij, vxn);
                    st_at(ovy, ij, vyn);
                    if constexpr (GEN) {
                        const int mo = mvi(O.mb, mco, i);
                        if (mo >= 0) {
                            const uint32_t o4 = (uint32_t)mo * (uint32_t)sizeof(T);
                            st_at(om[4], o4, m5);
                            st_at(om[5], o4, m6);
                            st_at(om[6], o4, m7);
                            st_at(om[7], o4, m8);
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 4; k++) WX[k] = WX[k + 1];
#pragma unroll
            for (int k = 0; k < 3; k++) WY[k] = WY[k + 1];
            LA[0] = LA[1];
            MU[0] = MU[1];
            RH[0] = RH[1];
        }
    };
    if (gen) walk(std::true_type{});
    else walk(std::false_type{});
}
template <typename T, int SUB>
__global__ __launch_bounds__(WAVE *NWAVES) void el_fwd_full(ElC<T> I, El<T> O, const T *__restrict__ rho,
                                                            const T *__restrict__ lam, const T *__restrict__ mu,
                                                            ElCoef<T> c, Rect R, StrmMap mp, int act)
{
    fwd_full_body<SUB, T>(I, O, rho, lam, mu, c, R, mp, act, (int)blockIdx.x);
}

// velocity phase, full grid: fw3/fw4 VJP on every cell
template <int SUB, typename T>
__device__ inline void adjv_full_body(const ElAdj<T> &a, const ElCoef<T> &c, const ElPar<T> &q,
                                      const StrmMap &mp, int bid)
{
    extern __shared__ __align__(16) unsigned char strm_lds[];
    T *__restrict__ gacc = reinterpret_cast<T *>(strm_lds);  // [wave][term][row][lane]
    int g, chunk, strip;
    if (!decode_strm(mp, g, chunk, strip, bid)) return;
    const int nw = (int)blockDim.x / WAVE;
    for (int k = threadIdx.x; k < nw * 2 * mp.ch * WAVE; k += blockDim.x) gacc[k] = T(0);
    __syncthreads();
    auto gslot = [&](int wv, int term, int r, int l) -> T & {
        return gacc[((wv * 2 + term) * mp.ch + r) * WAVE + l];
    };
    const int w = wave_id(), lane = threadIdx.x % WAVE;
    const int NXp = c.NX + 1, NYp = c.NY + 1;  // last padded row / column
    const int i0 = chunk * mp.ch, i1 = min(i0 + mp.ch, NXp + 1);
    const int jb = strip * SW_COLS - 2;
    const int j = jb + lane;
    const int jc = j < 0 ? 0 : (j > NYp ? NYp : j);  // loads of lanes beyond the grid stay inside it
    const bool out = lane >= 2 && lane < WAVE - 2 && j <= NYp;
    const bool gen = !(i0 - 2 >= a.R.iL && i1 + 2 <= a.R.iR && jb >= a.R.jL && jb + WAVE <= a.R.jR);
    const int s = g * a.Sg + w;
    if (w < a.Sg && s < a.act) {
        const int P = c.P;
        const T dt = c.dt;
        const T *__restrict__ vxa = a.A.f(VX, s);
        const T *__restrict__ vya = a.A.f(VY, s);
        T *__restrict__ axx = a.A.f(SXX, s);
        T *__restrict__ axy = a.A.f(SXY, s);
        T *__restrict__ ayy = a.A.f(SYY, s);
        const T *__restrict__ sxx = a.Sn.f(SXX, s);
        const T *__restrict__ sxy = a.Sn.f(SXY, s);
        const T *__restrict__ syy = a.Sn.f(SYY, s);
        const uint32_t pb = (uint32_t)P * (uint32_t)sizeof(T), jb4 = (uint32_t)jc * (uint32_t)sizeof(T);
        // DN: walk down the chunk (even chunks), else up (odd chunks), so the window rows two
        // adjacent chunks share are read by both wavefronts at about the same time
        auto walk = [&](auto genc, auto dnc) {
            constexpr bool GEN = decltype(genc)::value;
            constexpr bool DN = decltype(dnc)::value;
            constexpr int d = DN ? 1 : -1;
            const ProfRow<T> py = GEN ? prof_col(c, jc) : ProfRow<T>{};
            const FwPart fc = GEN ? fw_col<SUB>(c, j) : FwPart{true, true};
            const MvCol mc = mv_col(a.Sn.mb, jc);
            // raw loads of one window row, and of one output row
            struct Rw { T a, b, x, y, r, q, m5, m7; ProfRow<T> px; };
            struct Ro { T yy, m6, m8, gxx, gxy, gyy, r, q; };
            auto fetch_w = [&](int r, Rw &o) {
                const bool ok = !GEN || (r >= 0 && r <= NXp);  // uniform over the wavefront
                const uint32_t rj = (uint32_t)(ok ? r : 0) * pb + jb4;
                o.x = ok ? ld_at(sxx, rj) : T(0);
                o.y = ok ? ld_at(sxy, rj) : T(0);
                o.a = ok ? ld_at(vxa, rj) : T(0);
                o.b = ok ? ld_at(vya, rj) : T(0);
                o.r = ok ? ld_at(q.irho, rj) : T(0);
                o.q = ok ? ld_at(q.irb, rj) : T(0);
                // (cells whose profile is 0 keep the 0 of the clear: no test needed)
                o.m5 = GEN && ok ? ld_at(a.A.f(M5, s), rj) : T(0);
                o.m7 = GEN && ok ? ld_at(a.A.f(M7, s), rj) : T(0);
                o.px = GEN ? prof_row(c, r) : ProfRow<T>{};
            };
            auto fetch_o = [&](int i, Ro &o) {
                const uint32_t ij = (uint32_t)i * pb + jb4;
                o.yy = ld_at(syy, ij);
                o.r = ld_at(q.irho, ij);
                o.q = ld_at(q.irb, ij);
                o.m6 = GEN ? ld_at(a.A.f(M6, s), ij) : T(0);
                o.m8 = GEN ? ld_at(a.A.f(M8, s), ij) : T(0);
                o.gxx = out ? ld_at<T>(axx, ij) : T(0);
                o.gxy = out ? ld_at<T>(axy, ij) : T(0);
                o.gyy = out ? ld_at<T>(ayy, ij) : T(0);
            };
            // windows over rows i-2 .. i+2 (index 0 .. 4): t0, t2, e3 = g_vx dt / rho,
            // e4 = g_vy dt / rho-bar, the forward sxx, sxy; g5 = m5 + e3, g7 = m7 + e4
            T w0[5], w2[5], e3w[5], e4w[5], wx[5], wy[5], g5w[5], g7w[5];
            auto finish = [&](int r, const Rw &o, int k) {
                wx[k] = o.x;
                wy[k] = o.y;
                const T e3 = o.a * dt * o.r, e4 = o.b * dt * o.q;
                e3w[k] = e3;
                e4w[k] = e4;
                if constexpr (GEN) {
                    const FwPart fr = fw_row<SUB>(c, r);
                    const T g5 = o.m5 + e3, g7 = o.m7 + e4;
                    g5w[k] = g5;
                    g7w[k] = g7;
                    w0[k] = fr.B && fc.A ? e3 + o.px.af * g5 : T(0);
                    w2[k] = fr.A && fc.B ? e4 + o.px.ah * g7 : T(0);
                } else {
                    w0[k] = e3;
                    w2[k] = e4;
                }
            };
            const int ia = DN ? i0 : i1 - 1, nrow = i1 - i0;  // first output row, rows
#pragma unroll
            for (int k = 0; k < 4; k++) {  // the window rows but the one the first step takes
                const int kk = DN ? k : k + 1;
                Rw t;
                fetch_w(ia - 2 + kk, t);
                finish(ia - 2 + kk, t, kk);
            }
            Rw rn;
            Ro on;
            fetch_w(ia + 2 * d, rn);
            fetch_o(ia, on);
            int i = ia;
            for (int m = 0; m < nrow; m++, i += d) {
                const Rw rc = rn;
                const Ro oc = on;
                if (m + 1 < nrow) {
                    fetch_w(i + 3 * d, rn);
                    fetch_o(i + d, on);
                }
                finish(i + 2 * d, rc, DN ? 4 : 0);
                const FwPart fr = GEN ? fw_row<SUB>(c, i) : FwPart{true, true};
                const bool f3 = fr.B && fc.A, f4 = fr.A && fc.B;
                const T e3 = e3w[2], e4 = e4w[2];
                // row i, y-direction t-values
                const T g6 = oc.m6 + e3, g8 = oc.m8 + e4;
                const T t1 = GEN ? (f3 ? e3 + py.af * g6 : T(0)) : e3;
                const T t3 = GEN ? (f4 ? e4 + py.ah * g8 : T(0)) : e4;
                const T xy = wy[2], yy = oc.yy;
                const T t1m = dn1(t1), t1p = up1(t1), t1pp = up1(t1p);
                const T t3m = dn1(t3), t3mm = dn1(t3m), t3p = up1(t3);
                const T xym = dn1(xy), xymm = dn1(xym), xyp = up1(xy);
                const T yym = dn1(yy), yyp = up1(yy), yypp = up1(yyp);
                if (out) {
                    const uint32_t ij = (uint32_t)i * pb + jb4;
                    // parameter terms (fw3: rho; fw4: rho-bar), with the forward memory
                    // variables of level n outside the damping-free box
                    T m5 = T(0), m6 = T(0), m7 = T(0), m8 = T(0);
                    if constexpr (GEN) {
                        const int mi = mv_at(a.Sn.mb, mc, i);
                        if (mi >= 0) {
                            m5 = a.Sn.f(M5, s)[mi];
                            m6 = a.Sn.f(M6, s)[mi];
                            m7 = a.Sn.f(M7, s)[mi];
                            m8 = a.Sn.f(M8, s)[mi];
                        }
                    }
                    const T S3 = d4(wx[2], wx[1], wx[3], wx[0]) * c.ihx + m5 + d4(xy, xym, xyp, xymm) * c.ihy + m6;
                    const T S4 = d4(wy[3], wy[2], wy[4], wy[1]) * c.ihx + m7 + d4(yyp, yy, yypp, yym) * c.ihy + m8;
                    gslot(w, 0, i - i0, lane) = f3 ? -e3 * S3 * oc.r : T(0);
                    gslot(w, 1, i - i0, lane) = f4 ? -e4 * S4 * oc.q : T(0);
                    if constexpr (GEN) {
                        // memory-variable cotangents of level n-1 (those whose profile is
                        // non-zero)
                        const ProfRow<T> px = prof_row(c, i);
                        if (px.af != T(0)) st_at(a.mo[M5 - M1] + (int64_t)s * a.npad, ij, f3 ? px.bf * g5w[2] : T(0));
                        if (py.af != T(0)) st_at(a.mo[M6 - M1] + (int64_t)s * a.npad, ij, f3 ? py.bf * g6 : T(0));
                        if (px.ah != T(0)) st_at(a.mo[M7 - M1] + (int64_t)s * a.npad, ij, f4 ? px.bh * g7w[2] : T(0));
                        if (py.ah != T(0)) st_at(a.mo[M8 - M1] + (int64_t)s * a.npad, ij, f4 ? py.bh * g8 : T(0));
                    }
                    // sxx <- (D-x)^T t0 ; sxy <- (D-y)^T t1 + (D+x)^T t2 ; syy <- (D+y)^T t3
                    st_at(axx, ij, oc.gxx + d4(w0[2], w0[3], w0[1], w0[4]) * c.ihx);
                    st_at(axy, ij, oc.gxy + d4(t1, t1p, t1m, t1pp) * c.ihy + d4(w2[1], w2[2], w2[0], w2[3]) * c.ihx);
                    st_at(ayy, ij, oc.gyy + d4(t3m, t3, t3mm, t3p) * c.ihy);
                }
                if constexpr (DN) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        w0[k] = w0[k + 1];
                        w2[k] = w2[k + 1];
                        e3w[k] = e3w[k + 1];
                        e4w[k] = e4w[k + 1];
                        wx[k] = wx[k + 1];
                        wy[k] = wy[k + 1];
                        if constexpr (GEN) {
                            g5w[k] = g5w[k + 1];
                            g7w[k] = g7w[k + 1];
                        }
                    }
                } else {
#pragma unroll
                    for (int k = 4; k > 0; k--) {
                        w0[k] = w0[k - 1];
                        w2[k] = w2[k - 1];
                        e3w[k] = e3w[k - 1];
                        e4w[k] = e4w[k - 1];
                        wx[k] = wx[k - 1];
                        wy[k] = wy[k - 1];
                        if constexpr (GEN) {
                            g5w[k] = g5w[k - 1];
                            g7w[k] = g7w[k - 1];
                        }
                    }
                }
            }
        };
        const bool dn = (chunk & 1) == 0;
        if (gen) {
            if (dn) walk(std::true_type{}, std::true_type{});
            else walk(std::true_type{}, std::false_type{});
        } else {
            if (dn) walk(std::false_type{}, std::true_type{});
            else walk(std::false_type{}, std::false_type{});
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < (i1 - i0) * WAVE; k += blockDim.x) {
        const int r = k / WAVE, l = k % WAVE, jj = jb + l;
        if (l < 2 || l >= WAVE - 2 || jj > NYp) continue;
        const int64_t ij = (int64_t)(i0 + r) * c.P + jj;
        T s0 = T(0), s1 = T(0);
        for (int wv = 0; wv < nw; wv++) {
            s0 += gslot(wv, 0, r, l);
            s1 += gslot(wv, 1, r, l);
        }
        g_at(a, G_RHO, g, ij) += s0;
        g_at(a, G_RB, g, ij) += s1;
    }
}

// stress phase, full grid: fw1/fw2 VJP on every cell
template <int SUB, typename T>
__device__ inline void adjs_full_body(const ElAdj<T> &a, const ElCoef<T> &c, const ElPar<T> &q,
                                      const StrmMap &mp, int bid)
{
    extern __shared__ __align__(16) unsigned char strm_lds[];
    T *__restrict__ gacc = reinterpret_cast<T *>(strm_lds);  // [wave][term][row][lane]
    int g, chunk, strip;
    if (!decode_strm(mp, g, chunk, strip, bid)) return;
    const int nw = (int)blockDim.x / WAVE;
    for (int k = threadIdx.x; k < nw * 3 * mp.ch * WAVE; k += blockDim.x) gacc[k] = T(0);
    __syncthreads();
    auto gslot = [&](int wv, int term, int r, int l) -> T & {
        return gacc[((wv * 3 + term) * mp.ch + r) * WAVE + l];
    };
    const int w = wave_id(), lane = threadIdx.x % WAVE;
    const int NXp = c.NX + 1, NYp = c.NY + 1;
    const int i0 = chunk * mp.ch, i1 = min(i0 + mp.ch, NXp + 1);
    const int jb = strip * SW_COLS - 2;
    const int j = jb + lane;
    const int jc = j < 0 ? 0 : (j > NYp ? NYp : j);
    const bool out = lane >= 2 && lane < WAVE - 2 && j <= NYp;
    const bool gen = !(i0 - 2 >= a.R.iL && i1 + 2 <= a.R.iR && jb >= a.R.jL && jb + WAVE <= a.R.jR);
    const int s = g * a.Sg + w;
    if (w < a.Sg && s < a.act) {
        const int P = c.P;
        const T dt = c.dt;
        T *__restrict__ avx = a.A.f(VX, s);
        T *__restrict__ avy = a.A.f(VY, s);
        const T *__restrict__ axx = a.A.f(SXX, s);
        const T *__restrict__ ayy = a.A.f(SYY, s);
        const T *__restrict__ axy = a.A.f(SXY, s);
        const T *__restrict__ vx = a.Sm.f(VX, s);
        const T *__restrict__ vy = a.Sm.f(VY, s);
        const uint32_t pb = (uint32_t)P * (uint32_t)sizeof(T), jb4 = (uint32_t)jc * (uint32_t)sizeof(T);
        // DN: walk down the chunk (even chunks), else up (odd chunks), so the window rows two
        // adjacent chunks share are read by both wavefronts at about the same time
        auto walk = [&](auto genc, auto dnc) {
            constexpr bool GEN = decltype(genc)::value;
            constexpr bool DN = decltype(dnc)::value;
            constexpr int d = DN ? 1 : -1;
            const ProfRow<T> py = GEN ? prof_col(c, jc) : ProfRow<T>{};
            const FwPart fc = GEN ? fw_col<SUB>(c, j) : FwPart{true, true};
            const MvCol mc = mv_col(a.Sn.mb, jc);
            // raw loads of one window row (the stress cotangents, the averaged parameters, the
            // memory-variable cotangents of x-profile rows, vx and vy of level n-1), and of one
            // output row (its cotangents and parameters again, with those of y-profile
            // columns, and the velocity cotangents it updates)
            struct Rw { T u, v, xx, yy, xy, lb, mb, mb2, m1, m3; ProfRow<T> px; };
            struct Ro { T xx, yy, xy, lb, mb, mb2, m1, m2, m3, m4, gvx, gvy; };
            auto fetch_w = [&](int r, Rw &o) {
                const bool ok = !GEN || (r >= 0 && r <= NXp);  // uniform over the wavefront
                const uint32_t rj = (uint32_t)(ok ? r : 0) * pb + jb4;
                o.u = ok ? ld_at(vx, rj) : T(0);
                o.v = ok ? ld_at(vy, rj) : T(0);
                o.xx = ok ? ld_at(axx, rj) : T(0);
                o.yy = ok ? ld_at(ayy, rj) : T(0);
                o.xy = ok ? ld_at(axy, rj) : T(0);
                o.lb = ok ? ld_at(q.lb, rj) : T(0);
                o.mb = ok ? ld_at(q.mb, rj) : T(0);
                o.mb2 = ok ? ld_at(q.mb2, rj) : T(0);
                // (cells whose profile is 0 keep the 0 of the clear: no test needed)
                o.m1 = GEN && ok ? ld_at(a.A.f(M1, s), rj) : T(0);
                o.m3 = GEN && ok ? ld_at(a.A.f(M3, s), rj) : T(0);
                o.px = GEN ? prof_row(c, r) : ProfRow<T>{};
            };
            auto fetch_o = [&](int i, Ro &o) {
                const uint32_t ij = (uint32_t)i * pb + jb4;
                o.xx = ld_at(axx, ij);
                o.yy = ld_at(ayy, ij);
                o.xy = ld_at(axy, ij);
                o.lb = ld_at(q.lb, ij);
                o.mb = ld_at(q.mb, ij);
                o.mb2 = ld_at(q.mb2, ij);
                o.m1 = GEN ? ld_at(a.A.f(M1, s), ij) : T(0);
                o.m3 = GEN ? ld_at(a.A.f(M3, s), ij) : T(0);
                o.m2 = GEN ? ld_at(a.A.f(M2, s), ij) : T(0);
                o.m4 = GEN ? ld_at(a.A.f(M4, s), ij) : T(0);
                o.gvx = out ? ld_at<T>(avx, ij) : T(0);
                o.gvy = out ? ld_at<T>(avy, ij) : T(0);
            };
            // windows over rows i-2 .. i+2: t0, t2, and vx, vy of level n-1
            T w0[5], w2[5], wu[5], wv[5];
            auto finish = [&](int r, const Rw &o, int k) {
                wu[k] = o.u;
                wv[k] = o.v;
                const T gx = o.xx * dt, gy = o.yy * dt;
                const T lm = o.lb + T(2) * o.mb;
                const T gX = lm * gx + o.lb * gy;
                const T gs = o.xy * o.mb2 * dt;
                if constexpr (GEN) {
                    const FwPart fr = fw_row<SUB>(c, r);
                    w0[k] = fr.A && fc.A ? gX + o.px.ah * (o.m1 + gX) : T(0);
                    w2[k] = fr.B && fc.B ? gs + o.px.af * (o.m3 + gs) : T(0);
                } else {
                    w0[k] = gX;
                    w2[k] = gs;
                }
            };
            const int ia = DN ? i0 : i1 - 1, nrow = i1 - i0;  // first output row, rows
#pragma unroll
            for (int k = 0; k < 4; k++) {  // the window rows but the one the first step takes
                const int kk = DN ? k : k + 1;
                Rw t;
                fetch_w(ia - 2 + kk, t);
                finish(ia - 2 + kk, t, kk);
            }
            Rw rn;
            Ro on;
            fetch_w(ia + 2 * d, rn);
            fetch_o(ia, on);
            int i = ia;
            for (int m = 0; m < nrow; m++, i += d) {
                const Rw rc = rn;
                const Ro oc = on;
                if (m + 1 < nrow) {
                    fetch_w(i + 3 * d, rn);
                    fetch_o(i + d, on);
                }
                finish(i + 2 * d, rc, DN ? 4 : 0);
                const FwPart fr = GEN ? fw_row<SUB>(c, i) : FwPart{true, true};
                const bool f1 = fr.A && fc.A, f2 = fr.B && fc.B;
                // row i: the cotangent products of fw1 / fw2 and the y-direction t-values
                const T gx = oc.xx * dt, gy = oc.yy * dt;
                const T lm = oc.lb + T(2) * oc.mb;
                const T gX = lm * gx + oc.lb * gy, gY = oc.lb * gx + lm * gy;
                const T gs = oc.xy * oc.mb2 * dt;
                const T t1 = GEN ? (f1 ? gY + py.af * (oc.m2 + gY) : T(0)) : gY;
                const T t3 = GEN ? (f2 ? gs + py.ah * (oc.m4 + gs) : T(0)) : gs;
                const T u = wu[2], v = wv[2];
                const T t1m = dn1(t1), t1p = up1(t1), t1pp = up1(t1p);
                const T t3m = dn1(t3), t3mm = dn1(t3m), t3p = up1(t3);
                const T um = dn1(u), up = up1(u), upp = up1(up);
                const T vm = dn1(v), vmm = dn1(vm), vp = up1(v);
                if (out) {
                    const uint32_t ij = (uint32_t)i * pb + jb4;
                    // parameter terms: fw1 (lambda-bar, mu-bar) and fw2 (mu-bar), level n-1
                    // velocities, with the forward memory variables of level n outside the box
                    T m1 = T(0), m2 = T(0), m3 = T(0), m4 = T(0);
                    if constexpr (GEN) {
                        const int mi = mv_at(a.Sn.mb, mc, i);
                        if (mi >= 0) {
                            m1 = a.Sn.f(M1, s)[mi];
                            m2 = a.Sn.f(M2, s)[mi];
                            m3 = a.Sn.f(M3, s)[mi];
                            m4 = a.Sn.f(M4, s)[mi];
                        }
                    }
                    const T X = d4(wu[3], wu[2], wu[4], wu[1]) * c.ihx + m1;
                    const T Y = d4(v, vm, vp, vmm) * c.ihy + m2;
                    const T glm = X * gx + Y * gy;
                    gslot(w, 0, i - i0, lane) = f1 ? Y * gx + X * gy + glm : T(0);
                    gslot(w, 1, i - i0, lane) = f1 ? T(2) * glm : T(0);
                    const T dx = d4(wv[2], wv[1], wv[3], wv[0]) * c.ihx;
                    const T dy = d4(up, u, upp, um) * c.ihy;
                    gslot(w, 2, i - i0, lane) = f2 ? oc.xy * (dx + m3 + dy + m4) * dt : T(0);
                    if constexpr (GEN) {
                        // memory-variable cotangents of level n-1 (those whose profile is
                        // non-zero)
                        const ProfRow<T> px = prof_row(c, i);
                        if (px.ah != T(0)) st_at(a.mo[M1 - M1] + (int64_t)s * a.npad, ij, f1 ? px.bh * (oc.m1 + gX) : T(0));
                        if (py.af != T(0)) st_at(a.mo[M2 - M1] + (int64_t)s * a.npad, ij, f1 ? py.bf * (oc.m2 + gY) : T(0));
                        if (px.af != T(0)) st_at(a.mo[M3 - M1] + (int64_t)s * a.npad, ij, f2 ? px.bf * (oc.m3 + gs) : T(0));
                        if (py.ah != T(0)) st_at(a.mo[M4 - M1] + (int64_t)s * a.npad, ij, f2 ? py.bh * (oc.m4 + gs) : T(0));
                    }
                    // vx <- (D+x)^T t0 + (D+y)^T t3 ; vy <- (D-y)^T t1 + (D-x)^T t2
                    st_at(avx, ij, oc.gvx + d4(w0[1], w0[2], w0[0], w0[3]) * c.ihx + d4(t3m, t3, t3mm, t3p) * c.ihy);
                    st_at(avy, ij, oc.gvy + d4(t1, t1p, t1m, t1pp) * c.ihy + d4(w2[2], w2[3], w2[1], w2[4]) * c.ihx);
                }
                if constexpr (DN) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        w0[k] = w0[k + 1];
                        w2[k] = w2[k + 1];
                        wu[k] = wu[k + 1];
                        wv[k] = wv[k + 1];
                    }
                } else {
#pragma unroll
                    for (int k = 4; k > 0; k--) {
                        w0[k] = w0[k - 1];
                        w2[k] = w2[k - 1];
                        wu[k] = wu[k - 1];
                        wv[k] = wv[k - 1];
                    }
                }
            }
        };
        const bool dn = (chunk & 1) == 0;
        if (gen) {
            if (dn) walk(std::true_type{}, std::true_type{});
            else walk(std::true_type{}, std::false_type{});
        } else {
            if (dn) walk(std::false_type{}, std::true_type{});
            else walk(std::false_type{}, std::false_type{});
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < (i1 - i0) * WAVE; k += blockDim.x) {
        const int r = k / WAVE, l = k % WAVE, jj = jb + l;
        if (l < 2 || l >= WAVE - 2 || jj > NYp) continue;
        const int64_t ij = (int64_t)(i0 + r) * c.P + jj;
        T s0 = T(0), s1 = T(0), s2 = T(0);
        for (int wv = 0; wv < nw; wv++) {
            s0 += gslot(wv, 0, r, l);
            s1 += gslot(wv, 1, r, l);
            s2 += gslot(wv, 2, r, l);
        }
        g_at(a, G_LB, g, ij) += s0;
        g_at(a, G_MB1, g, ij) += s1;
        g_at(a, G_MB2, g, ij) += s2;
    }
}
template <typename T, int SUB>
__global__ __launch_bounds__(WAVE *SW_NW) void el_adjv_full(ElAdj<T> a, ElCoef<T> c, ElPar<T> q, StrmMap mp)
{
    adjv_full_body<SUB>(a, c, q, mp, (int)blockIdx.x);
}
template <typename T, int SUB>
__global__ __launch_bounds__(WAVE *SW_NW) void el_adjs_full(ElAdj<T> a, ElCoef<T> c, ElPar<T> q, StrmMap mp)
{
    adjs_full_body<SUB>(a, c, q, mp, (int)blockIdx.x);
}

// velocity phase: fw3/fw4 VJP on rectangle cells
template <typename T>
__device__ inline void adjv_strm_body(const ElAdj<T> &a, const ElCoef<T> &c, const ElPar<T> &q,
                                      const StrmMap &mp, int bid)
{
    extern __shared__ __align__(16) unsigned char strm_lds[];
    T *__restrict__ gacc = reinterpret_cast<T *>(strm_lds);  // [wave][term][row][lane]
    int g, chunk, strip;
    if (!decode_strm(mp, g, chunk, strip, bid)) return;
    const int nw = (int)blockDim.x / WAVE;
    for (int k = threadIdx.x; k < nw * 2 * mp.ch * WAVE; k += blockDim.x) gacc[k] = T(0);
    __syncthreads();
    auto gslot = [&](int wv, int term, int r, int l) -> T & {
        return gacc[((wv * 2 + term) * mp.ch + r) * WAVE + l];
    };
    const int w = wave_id(), lane = threadIdx.x % WAVE;
    const int i0 = a.R.iL + chunk * mp.ch, i1 = min(i0 + mp.ch, a.R.iR);
    const int jb = a.R.jL + strip * SW_COLS - 2;
    const int j = jb + lane;
    const int jc = j <= c.NY + 1 ? j : c.NY + 1;  // loads of lanes beyond the grid stay inside it
    const bool jin = j >= a.R.jL && j < a.R.jR;
    const bool jint = j >= 1 && j <= c.NY;
    const bool out = lane >= 2 && lane < WAVE - 2 && jin;
    const int s = g * a.Sg + w;
    if (w < a.Sg && s < a.act) {
        const int P = c.P;
        const int64_t ts = a.tstride;
        const T dt = c.dt;
        const T *__restrict__ vxa = a.A.f(VX, s);
        const T *__restrict__ vya = a.A.f(VY, s);
        T *__restrict__ axx = a.A.f(SXX, s);
        T *__restrict__ axy = a.A.f(SXY, s);
        T *__restrict__ ayy = a.A.f(SYY, s);
        const T *__restrict__ sxx = a.Sn.f(SXX, s);
        const T *__restrict__ sxy = a.Sn.f(SXY, s);
        const T *__restrict__ syy = a.Sn.f(SYY, s);
        const T *__restrict__ tb0 = a.Tb + (int64_t)s * a.npad;
        const T *__restrict__ tb1 = tb0 + ts;
        const T *__restrict__ tb2 = tb0 + 2 * ts;
        const T *__restrict__ tb3 = tb0 + 3 * ts;
        const uint32_t pb = (uint32_t)P * (uint32_t)sizeof(T), jb4 = (uint32_t)jc * (uint32_t)sizeof(T);
        // raw loads of one window row, and of one output row
        struct Rw { T a, b, x, y, r, q; };
        struct Ro { T yy, t1, t3, gxx, gxy, gyy; };
        auto fetch_w = [&](int r, Rw &o) {
            const uint32_t rj = (uint32_t)r * pb + jb4;
            const bool inr = r >= a.R.iL && r < a.R.iR;  // uniform over the wavefront
            const bool ld = inr || (jint && r >= 1 && r <= c.NX);
            o.x = ld_at(sxx, rj);
            o.y = ld_at(sxy, rj);
            o.a = ld ? ld_at(inr ? vxa : tb0, rj) : T(0);
            o.b = ld ? ld_at(inr ? vya : tb2, rj) : T(0);
            o.r = inr ? ld_at(q.irho, rj) : T(0);
            o.q = inr ? ld_at(q.irb, rj) : T(0);
        };
        auto fetch_o = [&](int i, Ro &o) {
            const uint32_t ij = (uint32_t)i * pb + jb4;
            o.yy = ld_at(syy, ij);
            o.t1 = !jin && jint ? ld_at(tb1, ij) : T(0);
            o.t3 = !jin && jint ? ld_at(tb3, ij) : T(0);
            o.gxx = out ? ld_at<T>(axx, ij) : T(0);
            o.gxy = out ? ld_at<T>(axy, ij) : T(0);
            o.gyy = out ? ld_at<T>(ayy, ij) : T(0);
        };
        // windows over rows i-2 .. i+2 (index 0 .. 4): t0, t2, sxx, sxy, 1/rho, 1/rho-bar;
        // t-values inside the rectangle from the cotangents, elsewhere the temporaries
        T w0[5], w2[5], wx[5], wy[5], wr[5], wb[5];
        auto finish = [&](int r, const Rw &o, int k) {
            wx[k] = o.x;
            wy[k] = o.y;
            wr[k] = o.r;
            wb[k] = o.q;
            const bool inr = r >= a.R.iL && r < a.R.iR;
            w0[k] = inr ? o.a * dt * o.r : o.a;
            w2[k] = inr ? o.b * dt * o.q : o.b;
        };
#pragma unroll
        for (int k = 0; k < 4; k++) {
            Rw t;
            fetch_w(i0 - 2 + k, t);
            finish(i0 - 2 + k, t, k);
        }
        Rw rn;
        Ro on;
        fetch_w(i0 + 2, rn);
        fetch_o(i0, on);
        for (int i = i0; i < i1; i++) {
            const Rw rc = rn;
            const Ro oc = on;
            if (i + 1 < i1) {
                fetch_w(i + 3, rn);
                fetch_o(i + 1, on);
            }
            finish(i + 2, rc, 4);
            // row i, y-direction inputs: t1, t3 (= t0, t2 inside the rectangle), sxy, syy
            const T t1 = jin ? w0[2] : oc.t1;
            const T t3 = jin ? w2[2] : oc.t3;
            const T xy = wy[2], yy = oc.yy;
            const T t1m = lane_at(t1, lane - 1), t1p = lane_at(t1, lane + 1), t1pp = lane_at(t1, lane + 2);
            const T t3mm = lane_at(t3, lane - 2), t3m = lane_at(t3, lane - 1), t3p = lane_at(t3, lane + 1);
            const T xymm = lane_at(xy, lane - 2), xym = lane_at(xy, lane - 1), xyp = lane_at(xy, lane + 1);
            const T yym = lane_at(yy, lane - 1), yyp = lane_at(yy, lane + 1), yypp = lane_at(yy, lane + 2);
            if (out) {
                const uint32_t ij = (uint32_t)i * pb + jb4;
                // parameter terms (fw3: rho; fw4: rho-bar)
                const T S3 = d4(wx[2], wx[1], wx[3], wx[0]) * c.ihx + d4(xy, xym, xyp, xymm) * c.ihy;
                const T S4 = d4(wy[3], wy[2], wy[4], wy[1]) * c.ihx + d4(yyp, yy, yypp, yym) * c.ihy;
                gslot(w, 0, i - i0, lane) = -w0[2] * S3 * wr[2];
                gslot(w, 1, i - i0, lane) = -w2[2] * S4 * wb[2];
                // sxx <- (D-x)^T t0 ; sxy <- (D-y)^T t1 + (D+x)^T t2 ; syy <- (D+y)^T t3
                st_at(axx, ij, oc.gxx + d4(w0[2], w0[3], w0[1], w0[4]) * c.ihx);
                st_at(axy, ij, oc.gxy + d4(t1, t1p, t1m, t1pp) * c.ihy + d4(w2[1], w2[2], w2[0], w2[3]) * c.ihx);
                st_at(ayy, ij, oc.gyy + d4(t3m, t3, t3mm, t3p) * c.ihy);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                w0[k] = w0[k + 1];
                w2[k] = w2[k + 1];
                wx[k] = wx[k + 1];
                wy[k] = wy[k + 1];
                wr[k] = wr[k + 1];
                wb[k] = wb[k + 1];
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < (i1 - i0) * WAVE; k += blockDim.x) {
        const int r = k / WAVE, l = k % WAVE, jj = jb + l;
        if (l < 2 || l >= WAVE - 2 || jj >= a.R.jR) continue;
        const int64_t ij = (int64_t)(i0 + r) * c.P + jj;
        T s0 = T(0), s1 = T(0);
        for (int wv = 0; wv < nw; wv++) {
            s0 += gslot(wv, 0, r, l);
            s1 += gslot(wv, 1, r, l);
        }
        g_at(a, G_RHO, g, ij) += s0;
        g_at(a, G_RB, g, ij) += s1;
    }
}

// stress phase: fw1/fw2 VJP on rectangle cells
template <typename T>
__device__ inline void adjs_strm_body(const ElAdj<T> &a, const ElCoef<T> &c, const ElPar<T> &q,
                                      const StrmMap &mp, int bid)
{
    extern __shared__ __align__(16) unsigned char strm_lds[];
    T *__restrict__ gacc = reinterpret_cast<T *>(strm_lds);  // [wave][term][row][lane]
    int g, chunk, strip;
    if (!decode_strm(mp, g, chunk, strip, bid)) return;
    const int nw = (int)blockDim.x / WAVE;
    for (int k = threadIdx.x; k < nw * 3 * mp.ch * WAVE; k += blockDim.x) gacc[k] = T(0);
    __syncthreads();
    auto gslot = [&](int wv, int term, int r, int l) -> T & {
        return gacc[((wv * 3 + term) * mp.ch + r) * WAVE + l];
    };
    const int w = wave_id(), lane = threadIdx.x % WAVE;
    const int i0 = a.R.iL + chunk * mp.ch, i1 = min(i0 + mp.ch, a.R.iR);
    const int jb = a.R.jL + strip * SW_COLS - 2;
    const int j = jb + lane;
    const int jc = j <= c.NY + 1 ? j : c.NY + 1;
    const bool jin = j >= a.R.jL && j < a.R.jR;
    const bool jint = j >= 1 && j <= c.NY;
    const bool out = lane >= 2 && lane < WAVE - 2 && jin;
    const int s = g * a.Sg + w;
    if (w < a.Sg && s < a.act) {
        const int P = c.P;
        const int64_t ts = a.tstride;
        const T dt = c.dt;
        T *__restrict__ avx = a.A.f(VX, s);
        T *__restrict__ avy = a.A.f(VY, s);
        const T *__restrict__ axx = a.A.f(SXX, s);
        const T *__restrict__ ayy = a.A.f(SYY, s);
        const T *__restrict__ axy = a.A.f(SXY, s);
        const T *__restrict__ vx = a.Sm.f(VX, s);
        const T *__restrict__ vy = a.Sm.f(VY, s);
        const T *__restrict__ tb0 = a.Tb + (int64_t)s * a.npad;
        const T *__restrict__ tb1 = tb0 + ts;
        const T *__restrict__ tb2 = tb0 + 2 * ts;
        const T *__restrict__ tb3 = tb0 + 3 * ts;
        const uint32_t pb = (uint32_t)P * (uint32_t)sizeof(T), jb4 = (uint32_t)jc * (uint32_t)sizeof(T);
        // raw loads of one window row (inside the rectangle's rows: the stress cotangents and the
        // parameters; outside: the temporaries t0..t3), and of one output row
        struct Rw { T u, v, e0, e1, e2, e3, e4, e5; };
        struct Ro { T t1, t3, gvx, gvy; };
        auto fetch_w = [&](int r, Rw &o) {
            const uint32_t rj = (uint32_t)r * pb + jb4;
            const bool inr = r >= a.R.iL && r < a.R.iR;  // uniform over the wavefront
            const bool ld = inr || (jint && r >= 1 && r <= c.NX);
            o.u = ld_at(vx, rj);
            o.v = ld_at(vy, rj);
            o.e0 = ld ? ld_at(inr ? axx : tb0, rj) : T(0);
            o.e1 = ld ? ld_at(inr ? ayy : tb1, rj) : T(0);
            o.e2 = ld ? ld_at(inr ? axy : tb2, rj) : T(0);
            o.e3 = ld ? ld_at(inr ? q.lb : tb3, rj) : T(0);
            o.e4 = inr ? ld_at(q.mb, rj) : T(0);
            o.e5 = inr ? ld_at(q.mb2, rj) : T(0);
        };
        auto fetch_o = [&](int i, Ro &o) {
            const uint32_t ij = (uint32_t)i * pb + jb4;
            o.t1 = !jin && jint ? ld_at(tb1, ij) : T(0);
            o.t3 = !jin && jint ? ld_at(tb3, ij) : T(0);
            o.gvx = out ? ld_at<T>(avx, ij) : T(0);
            o.gvy = out ? ld_at<T>(avy, ij) : T(0);
        };
        // windows over rows i-2 .. i+2: t0, t1, t2 (t3 = t2 inside the rectangle), the stress
        // cotangents sxx*dt, syy*dt, sxy of the row, and vx, vy of level n-1
        T w0[5], w1[5], w2[5], gx[5], gy[5], gs[5], wu[5], wv[5];
        auto finish = [&](int r, const Rw &o, int k) {
            wu[k] = o.u;
            wv[k] = o.v;
            if (r >= a.R.iL && r < a.R.iR) {
                gx[k] = o.e0 * dt;
                gy[k] = o.e1 * dt;
                gs[k] = o.e2;
                const T lb = o.e3, lm = lb + T(2) * o.e4;
                w0[k] = lm * gx[k] + lb * gy[k];
                w1[k] = lb * gx[k] + lm * gy[k];
                w2[k] = gs[k] * o.e5 * dt;
            } else {
                gx[k] = gy[k] = gs[k] = T(0);
                w0[k] = o.e0;
                w1[k] = o.e1;
                w2[k] = o.e2;  // t3 of such rows is never needed: only row i's t3 is read
            }
        };
#pragma unroll
        for (int k = 0; k < 4; k++) {
            Rw t;
            fetch_w(i0 - 2 + k, t);
            finish(i0 - 2 + k, t, k);
        }
        Rw rn;
        Ro on;
        fetch_w(i0 + 2, rn);
        fetch_o(i0, on);
        for (int i = i0; i < i1; i++) {
            const Rw rc = rn;
            const Ro oc = on;
            if (i + 1 < i1) {
                fetch_w(i + 3, rn);
                fetch_o(i + 1, on);
            }
            finish(i + 2, rc, 4);
            const T t1 = jin ? w1[2] : oc.t1, t3 = jin ? w2[2] : oc.t3, u = wu[2], v = wv[2];
            const T t1m = lane_at(t1, lane - 1), t1p = lane_at(t1, lane + 1), t1pp = lane_at(t1, lane + 2);
            const T t3mm = lane_at(t3, lane - 2), t3m = lane_at(t3, lane - 1), t3p = lane_at(t3, lane + 1);
            const T um = lane_at(u, lane - 1), up = lane_at(u, lane + 1), upp = lane_at(u, lane + 2);
            const T vmm = lane_at(v, lane - 2), vm = lane_at(v, lane - 1), vp = lane_at(v, lane + 1);
            if (out) {
                const uint32_t ij = (uint32_t)i * pb + jb4;
                // parameter terms: fw1 (lambda-bar, mu-bar) and fw2 (mu-bar), level n-1 velocities
                const T X = d4(wu[3], wu[2], wu[4], wu[1]) * c.ihx;
                const T Y = d4(v, vm, vp, vmm) * c.ihy;
                const T glm = X * gx[2] + Y * gy[2];
                gslot(w, 0, i - i0, lane) = Y * gx[2] + X * gy[2] + glm;
                gslot(w, 1, i - i0, lane) = T(2) * glm;
                const T dx = d4(wv[2], wv[1], wv[3], wv[0]) * c.ihx;
                const T dy = d4(up, u, upp, um) * c.ihy;
                gslot(w, 2, i - i0, lane) = gs[2] * (dx + dy) * dt;
                // vx <- (D+x)^T t0 + (D+y)^T t3 ; vy <- (D-y)^T t1 + (D-x)^T t2
                st_at(avx, ij, oc.gvx + d4(w0[1], w0[2], w0[0], w0[3]) * c.ihx + d4(t3m, t3, t3mm, t3p) * c.ihy);
                st_at(avy, ij, oc.gvy + d4(t1, t1p, t1m, t1pp) * c.ihy + d4(w2[2], w2[3], w2[1], w2[4]) * c.ihx);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                w0[k] = w0[k + 1];
                w1[k] = w1[k + 1];
                w2[k] = w2[k + 1];
                gx[k] = gx[k + 1];
                gy[k] = gy[k + 1];
                gs[k] = gs[k + 1];
                wu[k] = wu[k + 1];
                wv[k] = wv[k + 1];
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < (i1 - i0) * WAVE; k += blockDim.x) {
        const int r = k / WAVE, l = k % WAVE, jj = jb + l;
        if (l < 2 || l >= WAVE - 2 || jj >= a.R.jR) continue;
        const int64_t ij = (int64_t)(i0 + r) * c.P + jj;
        T s0 = T(0), s1 = T(0), s2 = T(0);
        for (int wv = 0; wv < nw; wv++) {
            s0 += gslot(wv, 0, r, l);
            s1 += gslot(wv, 1, r, l);
            s2 += gslot(wv, 2, r, l);
        }
        g_at(a, G_LB, g, ij) += s0;
        g_at(a, G_MB1, g, ij) += s1;
        g_at(a, G_MB2, g, ij) += s2;
    }
}

// The streaming rectangle alone, or fused with frame_b (independent cells, same phase):
// workgroups [0, nfb) take the frame items (nib item blocks per shot group, NWAVES items per
// workgroup), the rest the rectangle strips.  Both halves use WAVE * a.Sg threads; frame_b
// workgroups of fewer than NWAVES wavefronts are launched only when Sg == NWAVES.
template <typename T>
__global__ __launch_bounds__(WAVE *SW_NW) void el_adjv_strm(ElAdj<T> a, ElCoef<T> c, ElPar<T> q, StrmMap mp,
                                                            const int4 *__restrict__ items, int nitems,
                                                            int nib, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) adjv_frame_b_body(a, c, items, nitems, b % nib, b / nib);
    else adjv_strm_body(a, c, q, mp, b - nfb);
}
template <typename T>
__global__ __launch_bounds__(WAVE *SW_NW) void el_adjs_strm(ElAdj<T> a, ElCoef<T> c, ElPar<T> q, StrmMap mp,
                                                            const int4 *__restrict__ items, int nitems,
                                                            int nib, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) adjs_frame_b_body(a, c, items, nitems, b % nib, b / nib);
    else adjs_strm_body(a, c, q, mp, b - nfb);
}

// Shot-invariant inputs of the streaming kernels on every padded cell: 1/rho, 1/rho-bar
// (0 where the parameter is 0, the padding ring), lambda-bar and mu-bar of fw1
// ((i, j), (i+1, j) averages, Core.jl:109-110), mu-bar of fw2 ((i, j), (i, j+1), :141); the
// block form takes the parameters themselves.
template <typename T>
__global__ void el_adj_params(const T *__restrict__ rho, const T *__restrict__ rb,
                              const T *__restrict__ lam, const T *__restrict__ mu, T *__restrict__ irho,
                              T *__restrict__ irb, T *__restrict__ lb, T *__restrict__ mb,
                              T *__restrict__ mb2, int NX, int NY, int P, int sub)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= P) return;
    const int64_t ij = (int64_t)i * P + j;
    irho[ij] = rho[ij] != T(0) ? T(1) / rho[ij] : T(0);
    irb[ij] = rb[ij] != T(0) ? T(1) / rb[ij] : T(0);
    const bool dn = i <= NX, rt = j <= NY;  // (i+1, j) and (i, j+1) inside the padded grid
    lb[ij] = sub ? lam[ij] : (dn ? T(0.5) * (lam[ij + P] + lam[ij]) : T(0));
    mb[ij] = sub ? mu[ij] : (dn ? T(0.5) * (mu[ij + P] + mu[ij]) : T(0));
    mb2[ij] = sub ? mu[ij] : (rt ? T(0.5) * (mu[ij] + mu[ij + 1]) : T(0));
}

// rho-bar of fw4 at every cell of [1, NX-1] x [1, NY-1] (Core.jl:201), 0 elsewhere
template <typename T>
__global__ void el_rho_bar(const T *__restrict__ rho, T *__restrict__ rb, int NX, int NY, int P, int sub)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= P) return;
    const int64_t ij = (int64_t)i * P + j;
    if (sub) rb[ij] = (i >= 2 && i <= NX - 1 && j >= 2 && j <= NY - 1) ? rho[ij] : T(0);
    else
        rb[ij] = (i >= 1 && i <= NX - 1 && j >= 1 && j <= NY - 1)
                     ? T(0.25) * (rho[ij] + rho[ij + P] + rho[ij + P + 1] + rho[ij + 1])
                     : T(0);
}

// (rho, lambda, mu) gradients from the per-group accumulators, dense caller layout:
//   rho += G_RHO + (G_RB at (i,j), (i-1,j), (i-1,j-1), (i,j-1)) / 4
//   lambda += (G_LB at (i,j), (i-1,j)) / 2 ; mu += (G_MB1 at (i,j), (i-1,j)) / 2 + (G_MB2 at (i,j), (i,j-1)) / 2
template <typename T>
__global__ void el_grad_map(const T *__restrict__ G, int64_t npad, int64_t gstride, int ngroups,
                            T *__restrict__ grho, T *__restrict__ glam, T *__restrict__ gmu,
                            int NX, int NY, int P, int sub)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= NY + 2) return;
    auto at = [&](int k, int g, int ii, int jj) -> T {
        return ii >= 0 && jj >= 0 ? G[(int64_t)k * gstride + (int64_t)g * npad + (int64_t)ii * P + jj] : T(0);
    };
    T r = 0, l = 0, m = 0;
    for (int g = 0; g < ngroups; g++) {
        if (sub) {  // no averaging: each term belongs to its own cell
            r += at(G_RHO, g, i, j) + at(G_RB, g, i, j);
            l += at(G_LB, g, i, j);
            m += at(G_MB1, g, i, j) + at(G_MB2, g, i, j);
            continue;
        }
        r += at(G_RHO, g, i, j) +
             T(0.25) * (at(G_RB, g, i, j) + at(G_RB, g, i - 1, j) + at(G_RB, g, i - 1, j - 1) + at(G_RB, g, i, j - 1));
        l += T(0.5) * (at(G_LB, g, i, j) + at(G_LB, g, i - 1, j));
        m += T(0.5) * (at(G_MB1, g, i, j) + at(G_MB1, g, i - 1, j)) + T(0.5) * (at(G_MB2, g, i, j) + at(G_MB2, g, i, j - 1));
    }
    const int64_t d = (int64_t)i * (NY + 2) + j;
    grho[d] = r;
    glam[d] = l;
    gmu[d] = m;
}

// Start of an adjoint level, one launch: blocks [0, nxi) inject the trace residual of level
// n into the cotangents (unique (cell, type) per shot, blockIdx.y = shot), block nxi puts
// back the pre-injection stresses of the level's state (el_src_rcv mode 2).  The two touch
// different arrays.
struct SrcTab {
    const int64_t *us_ptr, *us_key, *us_eptr, *us_ent, *srcv_off, *nsrc, *rcv_ptr, *rcv_key, *nstep,
        *trace_off;
};
template <typename T>
__global__ __launch_bounds__(256) void el_adj_start(
    El<T> A, El<T> Sn, int n, const int64_t *__restrict__ ur_ptr, const int64_t *__restrict__ ur_key,
    const int64_t *__restrict__ ur_eptr, const int64_t *__restrict__ ur_ent,
    const int64_t *__restrict__ nstep, const T *__restrict__ resid,
    const int64_t *__restrict__ trace_off, int nxi, SrcTab st, const T *__restrict__ srcv,
    T *__restrict__ pre)
{
    const int s = blockIdx.y;
    if ((int)blockIdx.x >= nxi) {
        el_src_rcv_body(Sn, n, st.us_ptr, st.us_key, st.us_eptr, st.us_ent, srcv, st.srcv_off, st.nsrc,
                        st.rcv_ptr, st.rcv_key, st.nstep, (T *)nullptr, st.trace_off, 2, pre, s);
        return;
    }
    const int64_t c = ur_ptr[s] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ur_ptr[s + 1]) return;
    const int64_t key = ur_key[c], nt = nstep[s] + 1;
    const T *__restrict__ tr = resid + trace_off[s];
    T R = T(0);
    for (int64_t e = ur_eptr[c]; e < ur_eptr[c + 1]; e++) R += tr[ur_ent[e] * nt + n];
    A.f((int)(key & 7), s)[key >> 3] += R;
}

// dense cotangent of the 5 field histories at level n (full-history pullback)
template <typename T>
__global__ void el_adj_dense(El<T> A, ElC<T> D, int NX, int NY, int P)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int s = blockIdx.z;
    if (j >= NY + 2 || i >= NX + 2) return;
    const int64_t ij = (int64_t)i * P + j, d = (int64_t)i * (NY + 2) + j;  // D is dense
    for (int q = 0; q < 5; q++) A.f(q, s)[ij] += D.f(q, s)[d];
}

template <typename T>
__global__ void el_model_in(const T *__restrict__ m, T *__restrict__ o, int NX, int NY, int P)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= NY + 2) return;
    o[(int64_t)i * P + j] = m[(int64_t)i * (NY + 2) + j];
}


// ------------------------------- plan ---------------------------------------------------
struct ElPlanBase {
    virtual ~ElPlanBase() = default;
    ads_elastic_params p{};
    int dtype = ADS_F32;
    int64_t NX = 0, NY = 0, Pitch = 0, Npad = 0, Ndense = 0, NMAX = 0, nshots = 0;
    std::vector<int64_t> order, nstep_sorted, srcv_off_caller, trace_off_caller;
    int64_t srcv_elems = 0, trace_elems = 0;
    std::vector<int> active;
    double f_x = 0, f_y = 0;  // fractions of interior cells in the x / y CPML strips
    DevBuf d_prof;            // axf axh bxf bxh (NX each), ayf ayh byf byh (NY each)
    DevBuf d_us_ptr, d_us_eptr, d_us_ent, d_nsrc, d_srcv_off, d_ur_ptr, d_ur_eptr, d_ur_ent;
    DevBuf d_rcv_ptr, d_trace_off, d_nstep;
    DevBuf d_us_key[2], d_ur_key[2], d_rcv_key[2];  // [0] pitched, [1] dense; key = cell*8+type
    int64_t max_ur = 0;
    int64_t nkeys = 0;  // unique (cell, type) source keys over all shots
    // level slots and checkpoints keep the memory variables only outside this box, where some
    // CPML profile is non-zero (inside it they stay 0): nmv cells per shot and variable
    FrameBox mbox;
    int64_t nmv = 0;
    int64_t level_elems() const { return nshots * (5 * Npad + 8 * nmv); }  // one level, all shots
    int64_t field_elems(int q) const { return q < M1 ? Npad : nmv; }
    // interior rectangle (identity CPML, every update range) and frame items, per layout
    Rect rect[2]{};
    DevBuf d_items[2];
    int nitems[2] = {0, 0};
    DevBuf d_sitems;  // the pitched layout's frame as streaming items (FRS_ROWS x SW_FCOLS)
    int nsitems = 0;
    DevBuf d_ring[2];  // the rectangle's cells within two of its edge (adjoint t-values)
    int nring[2] = {0, 0};
    int rpt = 1, sg = 4;  // forward rectangle kernels: rows per wavefront, shots per workgroup
    // adjoint: shots per group (one wavefront each in the streaming rectangle kernels, whose
    // parameter terms are summed per group), streaming (1) or LDS-tiled (0) rectangle kernels,
    // and the streaming kernels' rows per workgroup
    int asg = NWAVES, strm = 2, strm_ch = 8;
    // forward: the fused streaming step (el_fwd_full) on the pitched layout, and its rows per
    // wavefront chunk (0: chosen per launch)
    int ffused = 1, fch = 0;
    // snapshot slots of the reversal schedule (< 0: one per segment start, the one-level scheme)
    int64_t snapshots = -1;
    int sub = 0;          // one MPIElastic subdomain (see ElCoef)
    int64_t PX = 0, PY = 0;  // CPML profile lengths (NX, NY; the block size in the subdomain form)
};

template <typename T> struct ElPlan : ElPlanBase {
    T dt_ = 0, ihx = 0, ihy = 0;

    ElCoef<T> coef(int P) const
    {
        const T *b = d_prof.as<T>();
        ElCoef<T> c;
        c.dt = dt_;
        c.ihx = ihx;
        c.ihy = ihy;
        c.axf = b;
        c.axh = b + PX;
        c.bxf = b + 2 * PX;
        c.bxh = b + 3 * PX;
        c.ayf = b + 4 * PX;
        c.ayh = b + 4 * PX + PY;
        c.byf = b + 4 * PX + 2 * PY;
        c.byh = b + 4 * PX + 3 * PY;
        c.prow = b + 4 * PX + 4 * PY;
        c.pcol = c.prow + 4 * (NX + 6);
        c.NX = (int)NX;
        c.NY = (int)NY;
        c.P = P;
        c.sub = sub;
        return c;
    }
    dim3 grid(int act) const { return dim3(cdiv(NY + 2, 64), cdiv(NX + 2, 4), act); }

    struct WS {
        int64_t K, NS, M, S;  // interval, slots, segments, snapshot slots
        T *seg, *ckpt, *adj, *tmp, *grad, *rho, *lam, *mu, *rb, *par, *resid;
        T *adjm;  // full-grid adjoint kernels: the second set of memory-variable cotangents
        T *pre;  // per level slot: the stress at each stress-type source key before injection
        double *loss, *part;
    };
    int64_t seg_count(int64_t K) const { return NMAX >= 1 ? (NMAX + K - 1) / K : 0; }
    // The full-grid adjoint kernels run at every level or at none: they keep the memory-variable
    // cotangents in two alternating sets, the other kernels in one set in place, so the choice
    // must not change with the number of active shots (their LDS is sized for the plan's group).
    bool full_adjoint() const
    {
        const size_t lds = (size_t)3 * asg * std::max(1, std::min(strm_ch, SW_CH_MAX)) * WAVE * sizeof(T);
        return strm == 2 && asg <= SW_NW && lds <= 65536 && (uint64_t)Npad * sizeof(T) < (uint64_t(1) << 32);
    }
    int64_t adj_groups() const { return (int64_t)cdiv(nshots, std::max(asg, 1)); }
    void layout(WS &w, Carve &cv, int64_t K) const { layout(w, cv, K, snapshots); }
    void layout(WS &w, Carve &cv, int64_t K, int64_t snaps) const
    {
        if (K < 1) K = 1;
        if (K > std::max<int64_t>(NMAX, 1)) K = std::max<int64_t>(NMAX, 1);
        w.K = K;
        w.M = seg_count(K);
        w.NS = std::min<int64_t>(K + 1, NMAX + 1);
        const int64_t one = std::max<int64_t>(w.M - 1, 0);
        w.S = snaps < 0 ? one : std::min(snaps, one);
        const size_t F = (size_t)nshots * Npad * sizeof(T), LV = (size_t)level_elems() * sizeof(T);
        w.seg = (T *)cv.take(LV * w.NS);
        w.ckpt = (T *)cv.take(LV * (size_t)w.S);  // slot q: a whole level of every shot
        w.adj = (T *)cv.take(F * NF);
        w.adjm = (T *)cv.take(full_adjoint() ? F * 8 : 0);
        w.tmp = (T *)cv.take(F * 4);
        w.grad = (T *)cv.take((size_t)G_COUNT * adj_groups() * Npad * sizeof(T));
        w.rb = (T *)cv.take((size_t)Npad * sizeof(T));
        w.par = (T *)cv.take((size_t)5 * Npad * sizeof(T));  // ElPar: irho, irb, lb, mb, mb2
        w.rho = (T *)cv.take((size_t)Npad * sizeof(T));
        w.lam = (T *)cv.take((size_t)Npad * sizeof(T));
        w.mu = (T *)cv.take((size_t)Npad * sizeof(T));
        w.resid = (T *)cv.take((size_t)std::max<int64_t>(trace_elems, 1) * sizeof(T));
        w.pre = (T *)cv.take((size_t)w.NS * (size_t)std::max<int64_t>(nkeys, 1) * sizeof(T));
        w.loss = (double *)cv.take(64);
        w.part = (double *)cv.take(MISFIT_BLOCKS * sizeof(double));
    }
    size_t ws_size(int64_t K) const { return ws_size(K, snapshots); }
    size_t ws_size(int64_t K, int64_t snaps) const
    {
        WS w;
        Carve cv(nullptr, 0);
        layout(w, cv, K, snaps);
        return cv.off + 256;
    }
    static constexpr double ADJ_COST = 2.5;  // an adjoint level costs about 2.5 forward levels
    int choose_schedule(size_t budget, int64_t *Kout, int64_t *Sout) const
    {
        double total = 0.0;
        for (int64_t n = 1; n <= NMAX; n++) total += active[n];
        auto ws = [&](int64_t K, int64_t S) { return ws_size(K, S); };
        auto work = [&](int64_t K, int64_t M) {
            std::vector<double> w((size_t)M, 0.0);
            for (int64_t m = 0; m < M; m++) {
                const int64_t a = m * K, b = std::min<int64_t>(a + K, NMAX);
                for (int64_t n = a + 1; n <= b; n++) w[m] += active[n];
            }
            return w;
        };
        return ads::choose_schedule(budget, NMAX, total, ADJ_COST, ws, work, *Kout, *Sout)
                   ? ADS_OK
                   : ADS_ERR_WORKSPACE;
    }
    int check_ws(WS &w, void *ws, size_t wsb, int64_t K)
    {
        Carve cv(ws, wsb);
        layout(w, cv, K);
        if (!ws || !cv.ok())
            ADS_FAIL(ADS_ERR_WORKSPACE, "workspace too small: need %zu bytes, got %zu", cv.off, wsb);
        return ADS_OK;
    }
    // [field][shot][Npad] block
    El<T> block(T *base) const
    {
        El<T> e;
        for (int q = 0; q < NF; q++) {
            e.p[q] = base + (size_t)q * nshots * Npad;
            e.st[q] = Npad;
        }
        return e;
    }
    // a level slot or checkpoint: the fields [field][shot][Npad], then the memory variables
    // [variable][shot][nmv] outside mbox
    El<T> level(T *base) const
    {
        El<T> e;
        T *q0 = base;
        for (int q = 0; q < NF; q++) {
            e.p[q] = q0;
            e.st[q] = field_elems(q);
            q0 += (size_t)nshots * field_elems(q);
        }
        e.mb = mbox;
        return e;
    }
    El<T> slot(const WS &w, int64_t n) const
    {
        return level(w.seg + (size_t)(n % w.NS) * level_elems());
    }
    // pre-injection stresses of the level in slot n (written by the forward step of level n;
    // the adjoint only reads levels that the last forward over their segment wrote)
    T *pre_slot(const WS &w, int64_t n) const { return w.pre + (size_t)(n % w.NS) * std::max<int64_t>(nkeys, 1); }

    static constexpr int VECP = sizeof(T) == 4 ? 4 : 2;  // pitched layout: 16-byte vectors
    // The rectangle of layout li, with the frame items in the same launch when `fuse` (the
    // pitched layout); otherwise the caller launches the frame.
    template <int RPT, int VEC>
    void launch_rect_r(int phase, const Rect &R, const ElC<T> &I, const El<T> &O, const T *rho,
                       const T *lam, const T *mu, const ElCoef<T> &c, int act, int Sg, int li,
                       bool fuse, hipStream_t st)
    {
        TileMap m;
        m.ntj = (int)cdiv(R.jR - R.jL, WAVE * VEC);
        m.nstrips = (int)cdiv(R.iR - R.iL, NWAVES * RPT);
        m.ngroups = (int)cdiv(act, Sg);
        m.nb = m.ntj * m.nstrips * m.ngroups;
        const unsigned nb = cdiv(m.nb, NXCD) * NXCD;
        if (fuse) {  // the frame as streaming items (the pitched layout's rectangle)
            const int nib = (int)cdiv(nsitems, NWAVES), nfb = nib * act;
            const int4 *it = d_sitems.template as<int4>();
            if (phase == 0)
                el_stress_level<T, RPT, VEC><<<nb + nfb, WAVE * NWAVES, 0, st>>>(I, O, lam, mu, c, R, m, act, Sg,
                                                                                it, nsitems, nib, nfb);
            else
                el_vel_level<T, RPT, VEC><<<nb + nfb, WAVE * NWAVES, 0, st>>>(I, O, rho, c, R, m, act, Sg,
                                                                             it, nsitems, nib, nfb);
            return;
        }
        if (phase == 0) el_stress_rect<T, RPT, VEC><<<nb, WAVE * NWAVES, 0, st>>>(I, O, lam, mu, c, R, m, act, Sg);
        else el_vel_rect<T, RPT, VEC><<<nb, WAVE * NWAVES, 0, st>>>(I, O, rho, c, R, m, act, Sg);
    }
    template <int VEC>
    void launch_rect(int phase, const Rect &R, const ElC<T> &I, const El<T> &O, const T *rho,
                     const T *lam, const T *mu, const ElCoef<T> &c, int act, int Sg, int li, bool fuse,
                     hipStream_t st)
    {
        if (rpt == 2) launch_rect_r<2, VEC>(phase, R, I, O, rho, lam, mu, c, act, Sg, li, fuse, st);
        else launch_rect_r<1, VEC>(phase, R, I, O, rho, lam, mu, c, act, Sg, li, fuse, st);
    }

    int step(const El<T> &I, const El<T> &O, int n, int act, const T *rho, const T *lam,
             const T *mu, const T *srcv, T *traces, int record, int P, int li, hipStream_t st,
             T *pre = nullptr)
    {
        if (act <= 0) return ADS_OK;
        const ElCoef<T> c = coef(P);
        const double cells = (double)act * (double)PX * (double)PY;  // interior (or block) cells
        // algorithmic bytes, the whole-step count w (13 + 8 f_x + 8 f_y) with each phase charged
        // for the fields it updates: stress phase sigma read + write (6), lam, mu (2) and m1..m4
        // in their strips; velocity phase v read + write (4), rho (1) and m5..m8.  The reads of
        // the other phase's fields are the stencil neighbours a fused step would reuse on chip.
        const Rect &R = rect[li];
        const bool rr = R.iR > R.iL && R.jR > R.jL;
        const int Sg = std::min(sg, act);
        const dim3 fg(cdiv(nitems[li], NWAVES), 1, act);
        const bool fuse = rr && li == 0 && nitems[li] > 0;  // rectangle + frame in one launch
        if (li == 0 && ffused && (uint64_t)Npad * sizeof(T) < (uint64_t(1) << 32)) {
            // one launch for both phases (the whole step's bytes, charged to the stress class)
            // Taller chunks re-read fewer halo rows (8 rows: 1.19x the step's compulsory bytes,
            // 32 rows: 1.0x at 600^2 x 16 shots) but give fewer wavefronts, each walking its
            // rows one after another: below about 12 per SIMD the walk's latency is exposed
            // (32 rows there: 130 against 104 us per level).  The tallest of 8, 16 and 32 rows
            // that keeps that many wavefronts.
            StrmMap mp;
            mp.nstrip = (int)cdiv(NY + 2, FS_COLS);
            mp.ch = fch;
            if (mp.ch <= 0) {
                mp.ch = 8;
                for (int h = 16; h <= 32; h *= 2)
                    if ((int64_t)mp.nstrip * cdiv(NX + 2, h) * act >= 12288) mp.ch = h;
            }
            mp.nchunk = (int)cdiv(NX + 2, mp.ch);
            mp.ngroup = (int)cdiv(act, NWAVES);
            mp.nb = mp.nstrip * mp.nchunk * mp.ngroup;
            const unsigned nb = (unsigned)(cdiv(mp.nb, NXCD) * NXCD);
            const int pr = prof_start(ADS_PROF_EL_FWD_STRESS, st);
            if (sub) el_fwd_full<T, 1><<<nb, WAVE * NWAVES, 0, st>>>(cst(I), O, rho, lam, mu, c, R, mp, act);
            else el_fwd_full<T, 0><<<nb, WAVE * NWAVES, 0, st>>>(cst(I), O, rho, lam, mu, c, R, mp, act);
            prof_stop(pr, st, cells * sizeof(T) * (13.0 + 8.0 * f_x + 8.0 * f_y));
        } else {
        int pr = prof_start(ADS_PROF_EL_FWD_STRESS, st);
        if (rr) {
            if (li == 0) launch_rect<VECP>(0, R, cst(I), O, rho, lam, mu, c, act, Sg, li, fuse, st);
            else launch_rect<1>(0, R, cst(I), O, rho, lam, mu, c, act, Sg, li, false, st);
        }
        if (nitems[li] && !fuse)
            el_stress_frame<T><<<fg, WAVE * NWAVES, 0, st>>>(cst(I), O, lam, mu, c,
                                                              d_items[li].template as<int4>(), nitems[li]);
        prof_stop(pr, st, cells * sizeof(T) * (8.0 + 4.0 * f_x + 4.0 * f_y));
        pr = prof_start(ADS_PROF_EL_FWD_VEL, st);
        if (rr) {
            if (li == 0) launch_rect<VECP>(1, R, cst(I), O, rho, lam, mu, c, act, Sg, li, fuse, st);
            else launch_rect<1>(1, R, cst(I), O, rho, lam, mu, c, act, Sg, li, false, st);
        }
        if (nitems[li] && !fuse)
            el_vel_frame<T><<<fg, WAVE * NWAVES, 0, st>>>(cst(I), O, rho, c,
                                                           d_items[li].template as<int4>(), nitems[li]);
        prof_stop(pr, st, cells * sizeof(T) * (5.0 + 4.0 * f_x + 4.0 * f_y));
        }
        el_src_rcv<T><<<act, 256, 0, st>>>(O, n, d_us_ptr.as<int64_t>(), d_us_key[li].template as<int64_t>(),
                                           d_us_eptr.as<int64_t>(), d_us_ent.as<int64_t>(), srcv,
                                           d_srcv_off.as<int64_t>(), d_nsrc.as<int64_t>(),
                                           d_rcv_ptr.as<int64_t>(), d_rcv_key[li].template as<int64_t>(),
                                           d_nstep.as<int64_t>(), traces,
                                           d_trace_off.as<int64_t>(), record ? 1 : 0, pre);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int prepare_models(T *dr, T *dl, T *dm, const void *rho, const void *lam, const void *mu,
                       hipStream_t st)
    {
        dim3 g(cdiv(NY + 2, 256), NX + 2);
        el_model_in<T><<<g, 256, 0, st>>>((const T *)rho, dr, (int)NX, (int)NY, (int)Pitch);
        el_model_in<T><<<g, 256, 0, st>>>((const T *)lam, dl, (int)NX, (int)NY, (int)Pitch);
        el_model_in<T><<<g, 256, 0, st>>>((const T *)mu, dm, (int)NX, (int)NY, (int)Pitch);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int run_forward(const WS &w, int64_t lo, int64_t hi, const T *srcv, T *traces, int record,
                    hipStream_t st)
    {
        for (int64_t n = lo + 1; n <= hi; n++)
            ADS_TRY(step(slot(w, n - 1), slot(w, n), (int)n, active[n], w.rho, w.lam, w.mu, srcv,
                         traces, record, (int)Pitch, 0, st, pre_slot(w, n)));
        return ADS_OK;
    }

    int copy_state(const El<T> &dst, const El<T> &src, int act, hipStream_t st)
    {
        for (int q = 0; q < NF; q++)
            ADS_HIP(hipMemcpyAsync(dst.p[q], src.p[q], (size_t)act * field_elems(q) * sizeof(T),
                                   hipMemcpyDeviceToDevice, st));
        return ADS_OK;
    }
    El<T> ckpt(const WS &w, int64_t q) const  // snapshot slot q
    {
        return level(w.ckpt + (size_t)q * level_elems());
    }

    int forward(const void *rho, const void *lam, const void *mu, const void *srcv, void *traces,
                void *ws, size_t wsb, int64_t K, hipStream_t st)
    {
        WS w;
        ADS_TRY(check_ws(w, ws, wsb, K));
        ADS_TRY(prepare_models(w.rho, w.lam, w.mu, rho, lam, mu, st));
        // Level 0 is zero everywhere.  The CPML memory variables and their cotangents are read
        // and written on frame cells only (frame kernels); rectangle cells never touch them, so
        // the other slots need no clearing.
        ADS_HIP(hipMemsetAsync(slot(w, 0).p[0], 0, (size_t)level_elems() * sizeof(T), st));
        if (traces && trace_elems) ADS_HIP(hipMemsetAsync(traces, 0, trace_elems * sizeof(T), st));
        // the reversal schedule's first descent: the states it keeps, in level order
        Reversal rv(w.M, w.S);
        const auto path = rv.first_path();
        size_t pi = 0;
        for (int64_t m = 0; m < w.M; m++) {
            int64_t a = m * w.K, b = std::min<int64_t>(a + w.K, NMAX);
            if (pi < path.size() && path[pi].first == m)
                ADS_TRY(copy_state(ckpt(w, path[pi++].second), slot(w, a), active[std::min<int64_t>(a + 1, NMAX)], st));
            ADS_TRY(run_forward(w, a, b, (const T *)srcv, (T *)traces, traces != nullptr, st));
        }
        return ADS_OK;
    }

    // one reverse level: cotangents of level n -> level n-1.  Sn: state of level n, Sm: n-1.
    int adj_level(const El<T> &A, const El<T> &Sn, const El<T> &Sm, int n, int act,
                  const T *resid, const ElC<T> *dense, const WS &w, const T *srcv, int P, int li,
                  hipStream_t st)
    {
        if (act <= 0) return ADS_OK;
        const ElCoef<T> c = coef(P);
        {
            const int nxi = resid && max_ur > 0 ? (int)cdiv(max_ur, 256) : 0;
            const SrcTab tab{d_us_ptr.as<int64_t>(),    d_us_key[li].template as<int64_t>(),
                             d_us_eptr.as<int64_t>(),   d_us_ent.as<int64_t>(),
                             d_srcv_off.as<int64_t>(),  d_nsrc.as<int64_t>(),
                             d_rcv_ptr.as<int64_t>(),   d_rcv_key[li].template as<int64_t>(),
                             d_nstep.as<int64_t>(),     d_trace_off.as<int64_t>()};
            el_adj_start<T><<<dim3(nxi + 1, act), 256, 0, st>>>(
                A, Sn, n, d_ur_ptr.as<int64_t>(), d_ur_key[li].template as<int64_t>(), d_ur_eptr.as<int64_t>(),
                d_ur_ent.as<int64_t>(), d_nstep.as<int64_t>(), resid, d_trace_off.as<int64_t>(), nxi, tab, srcv,
                pre_slot(w, n));
        }
        if (dense) el_adj_dense<T><<<grid(act), 256, 0, st>>>(A, *dense, (int)NX, (int)NY, P);
        ElAdj<T> a;
        a.A = A;
        a.Sn = cst(Sn);
        a.Sm = cst(Sm);
        a.Tb = w.tmp;
        a.G = w.grad;
        a.npad = Npad;
        a.tstride = (int64_t)nshots * Npad;
        a.gstride = adj_groups() * Npad;
        a.rho = w.rho;
        a.lam = w.lam;
        a.mu = w.mu;
        a.rb = w.rb;
        a.R = rect[li];
        a.act = act;
        a.Sg = std::min(asg, act);
        const int ng = (int)cdiv(act, a.Sg);
        const dim3 fg(cdiv(nitems[li], NWAVES), ng);
        const int4 *items = d_items[li].template as<int4>();
        const bool rr = a.R.iR > a.R.iL && a.R.jR > a.R.jL;
        TileMap m{};
        if (rr) {
            m.ntj = (int)cdiv(a.R.jR - a.R.jL, LT_J);
            m.nstrips = (int)cdiv(a.R.iR - a.R.iL, LT_I);
            m.ngroups = ng;
            m.nb = m.ntj * m.nstrips * m.ngroups;
        }
        const unsigned rg = cdiv(m.nb, NXCD) * NXCD;
        const double cells = (double)act * (double)PX * (double)PY;  // interior (or block) cells
        // algorithmic bytes, the whole-step count w (24 + 8 f_x + 8 f_y) with each phase charged
        // for what it updates: velocity phase the stress cotangents read + write (6), the
        // stresses (3), rho (1) and its gradient read + write (2); stress phase the velocity
        // cotangents read + write (4), the velocities (2), lambda, mu (2) and their gradients
        // read + write (4); each phase 4 CPML memory cotangents in their strips
        const ElPar<T> par{w.par, w.par + Npad, w.par + 2 * Npad, w.par + 3 * Npad, w.par + 4 * Npad};
        // streaming kernels: their LDS (3 terms x shots x rows x WAVE values) must fit in 64 KiB,
        // else the LDS-tiled kernels run
        const size_t strm_lds = (size_t)3 * a.Sg * std::max(1, std::min(strm_ch, SW_CH_MAX)) * WAVE * sizeof(T);
        // (the streaming kernels address one field of one shot with 32-bit byte offsets)
        const bool off32 = (uint64_t)Npad * sizeof(T) < (uint64_t(1) << 32);
        if (full_adjoint() && li == 0) {
            // full grid, one launch per phase.  Level n reads the memory-variable cotangents
            // that level n+1 wrote: the two sets alternate with the parity of n.
            El<T> Ar = A;
            for (int f = M1; f <= M8; f++) {
                T *other = w.adjm + (size_t)(f - M1) * nshots * Npad;
                if (n & 1) {
                    Ar.p[f] = other;
                    a.mo[f - M1] = A.p[f];
                } else {
                    a.mo[f - M1] = other;
                }
            }
            a.A = Ar;
            StrmMap sm;
            sm.ch = std::max(1, std::min(strm_ch, SW_CH_MAX));
            sm.nstrip = (int)cdiv(NY + 2, SW_COLS);
            sm.nchunk = (int)cdiv(NX + 2, sm.ch);
            sm.ngroup = ng;
            sm.nb = sm.nstrip * sm.nchunk * sm.ngroup;
            const unsigned sgrid = cdiv(sm.nb, NXCD) * NXCD;
            const unsigned sblock = (unsigned)(WAVE * a.Sg);
            const size_t lds = (size_t)a.Sg * sm.ch * WAVE * sizeof(T);  // x terms
            int pr = prof_start(ADS_PROF_EL_ADJ_VEL, st);
            if (sub) el_adjv_full<T, 1><<<sgrid, sblock, 2 * lds, st>>>(a, c, par, sm);
            else el_adjv_full<T, 0><<<sgrid, sblock, 2 * lds, st>>>(a, c, par, sm);
            prof_stop(pr, st, cells * sizeof(T) * (12.0 + 4.0 * (f_x + f_y)));
            pr = prof_start(ADS_PROF_EL_ADJ_STRESS, st);
            if (sub) el_adjs_full<T, 1><<<sgrid, sblock, 3 * lds, st>>>(a, c, par, sm);
            else el_adjs_full<T, 0><<<sgrid, sblock, 3 * lds, st>>>(a, c, par, sm);
            prof_stop(pr, st, cells * sizeof(T) * (12.0 + 4.0 * (f_x + f_y)));
            ADS_LAUNCH_CHECK();
            return ADS_OK;
        }
        if (rr && strm && li == 0 && a.Sg <= SW_NW && strm_lds <= 65536 && off32) {
            StrmMap sm;
            sm.ch = std::max(1, std::min(strm_ch, SW_CH_MAX));
            sm.nstrip = (int)cdiv(a.R.jR - a.R.jL, SW_COLS);
            sm.nchunk = (int)cdiv(a.R.iR - a.R.iL, sm.ch);
            sm.ngroup = ng;
            sm.nb = sm.nstrip * sm.nchunk * sm.ngroup;
            const unsigned sgrid = cdiv(sm.nb, NXCD) * NXCD;
            const unsigned sblock = (unsigned)(WAVE * a.Sg);
            const int nib = (int)cdiv(nitems[li], NWAVES), nfa = nib * ng;
            const int nrb = (int)cdiv(nring[li], NWAVES);
            const int4 *ring = d_ring[li].template as<int4>();
            // frame_b rides in the streaming launch when its workgroups have NWAVES wavefronts
            const bool fuse = nitems[li] > 0 && a.Sg == NWAVES;
            const int nfb = fuse ? nfa : 0;
            int pr = prof_start(ADS_PROF_EL_ADJ_VEL, st);
            if (nfa) el_adjv_a<T><<<nfa, WAVE * NWAVES, 0, st>>>(a, c, items, nitems[li], std::max(nib, 1), ring,
                                                                 0, 1, nfa);
            if (nitems[li] && !fuse) el_adjv_frame_b<T><<<fg, WAVE * NWAVES, 0, st>>>(a, c, items, nitems[li]);
            const size_t lds = (size_t)a.Sg * sm.ch * WAVE * sizeof(T);  // x terms
            el_adjv_strm<T><<<(unsigned)nfb + sgrid, sblock, 2 * lds, st>>>(a, c, par, sm, items, nitems[li],
                                                                              std::max(nib, 1), nfb);
            prof_stop(pr, st, cells * sizeof(T) * (12.0 + 4.0 * (f_x + f_y)));
            pr = prof_start(ADS_PROF_EL_ADJ_STRESS, st);
            const unsigned ga = (unsigned)(nfa + nrb * ng);
            if (ga) el_adjs_a<T><<<ga, WAVE * NWAVES, 0, st>>>(a, c, items, nitems[li], std::max(nib, 1), ring,
                                                               nring[li], std::max(nrb, 1), nfa);
            if (nitems[li] && !fuse) el_adjs_frame_b<T><<<fg, WAVE * NWAVES, 0, st>>>(a, c, items, nitems[li]);
            el_adjs_strm<T><<<(unsigned)nfb + sgrid, sblock, 3 * lds, st>>>(a, c, par, sm, items, nitems[li],
                                                                              std::max(nib, 1), nfb);
            prof_stop(pr, st, cells * sizeof(T) * (12.0 + 4.0 * (f_x + f_y)));
            ADS_LAUNCH_CHECK();
            return ADS_OK;
        }
        int pr = prof_start(ADS_PROF_EL_ADJ_VEL, st);
        const int nib = (int)cdiv(nitems[li], NWAVES), nfb = nib * ng;
        // velocity phase: frame_a alone (its gathers recompute rectangle t-values, two loads
        // each); stress phase: frame_a with the rectangle's ring, whose t-values cost five
        // parameter loads each to recompute, so frame_b reads them from the temporaries
        const int nrb = (int)cdiv(nring[li], NWAVES), nfa = nib * ng;
        const unsigned ga = (unsigned)(nfa + nrb * ng);
        const int4 *ring = d_ring[li].template as<int4>();
        if (nfa) el_adjv_a<T><<<nfa, WAVE * NWAVES, 0, st>>>(a, c, items, nitems[li], std::max(nib, 1), ring,
                                                             0, 1, nfa);
        if (rr && nitems[li])
            el_adjv_rb<T><<<(unsigned)nfb + rg, WAVE * NWAVES, 0, st>>>(a, c, m, items, nitems[li], nib, nfb);
        else if (rr) el_adjv_rect<T><<<rg, WAVE * NWAVES, 0, st>>>(a, c, m);
        else if (nitems[li]) el_adjv_frame_b<T><<<fg, WAVE * NWAVES, 0, st>>>(a, c, items, nitems[li]);
        prof_stop(pr, st, cells * sizeof(T) * (12.0 + 4.0 * (f_x + f_y)));
        pr = prof_start(ADS_PROF_EL_ADJ_STRESS, st);
        if (ga) el_adjs_a<T><<<ga, WAVE * NWAVES, 0, st>>>(a, c, items, nitems[li], std::max(nib, 1), ring,
                                                           nring[li], std::max(nrb, 1), nfa);
        if (rr && nitems[li])
            el_adjs_rb<T><<<(unsigned)nfb + rg, WAVE * NWAVES, 0, st>>>(a, c, m, items, nitems[li], nib, nfb);
        else if (rr) el_adjs_rect<T><<<rg, WAVE * NWAVES, 0, st>>>(a, c, m);
        else if (nitems[li]) el_adjs_frame_b<T><<<fg, WAVE * NWAVES, 0, st>>>(a, c, items, nitems[li]);
        prof_stop(pr, st, cells * sizeof(T) * (12.0 + 4.0 * (f_x + f_y)));
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int backward(const void *srcv, const T *resid, const void *dense_hist, void *grho, void *glam,
                 void *gmu, void *ws, size_t wsb, int64_t K, hipStream_t st)
    {
        WS w;
        ADS_TRY(check_ws(w, ws, wsb, K));
        const size_t F = (size_t)nshots * Npad * sizeof(T);
        ADS_HIP(hipMemsetAsync(w.adj, 0, F * NF, st));
        if (full_adjoint()) ADS_HIP(hipMemsetAsync(w.adjm, 0, F * 8, st));
        ADS_HIP(hipMemsetAsync(w.grad, 0, (size_t)G_COUNT * adj_groups() * Npad * sizeof(T), st));
        ADS_HIP(hipMemsetAsync(w.tmp, 0, F * 4, st));
        el_rho_bar<T><<<dim3(cdiv(Pitch, 256), NX + 2), 256, 0, st>>>(w.rho, w.rb, (int)NX, (int)NY, (int)Pitch, sub);
        el_adj_params<T><<<dim3(cdiv(Pitch, 256), NX + 2), 256, 0, st>>>(
            w.rho, w.rb, w.lam, w.mu, w.par, w.par + Npad, w.par + 2 * Npad, w.par + 3 * Npad,
            w.par + 4 * Npad, (int)NX, (int)NY, (int)Pitch, sub);
        const T *sv = (const T *)srcv;
        const El<T> A = block(w.adj);
        const int64_t L = NMAX + 1;
        struct Ops {
            ElPlan *pl;
            const WS &w;
            const El<T> &A;
            const T *sv, *resid;
            const void *dense_hist;
            int64_t L;
            hipStream_t st;
            int act_at(int64_t m) const { return pl->active[std::min<int64_t>(m * w.K + 1, pl->NMAX)]; }
            int restore(int64_t m, int64_t q)
            {
                const int act = act_at(m);
                if (q < 0) {  // level 0: zero
                    El<T> z = pl->slot(w, 0);
                    for (int f = 0; f < NF; f++)
                        ADS_HIP(hipMemsetAsync(z.p[f], 0, (size_t)act * pl->field_elems(f) * sizeof(T), st));
                    return ADS_OK;
                }
                return pl->copy_state(pl->slot(w, m * w.K), pl->ckpt(w, q), act, st);
            }
            int save(int64_t m, int64_t q) { return pl->copy_state(pl->ckpt(w, q), pl->slot(w, m * w.K), act_at(m), st); }
            int forward(int64_t m0, int64_t m1)
            {
                for (int64_t m = m0; m < m1; m++) {
                    const int64_t a = m * w.K, b = std::min<int64_t>(a + w.K, pl->NMAX);
                    ADS_TRY(pl->run_forward(w, a, b, sv, nullptr, 0, st));
                }
                return ADS_OK;
            }
            int adjoint(int64_t m)
            {
                const int64_t a = m * w.K, b = std::min<int64_t>(a + w.K, pl->NMAX);
                for (int64_t n = b; n > a; n--) {
                    ElC<T> D, *Dp = nullptr;
                    if (dense_hist) {  // [shot][field][level][Ndense]
                        for (int q = 0; q < 5; q++) {
                            D.p[q] = (const T *)dense_hist + ((size_t)q * L + n) * pl->Ndense;
                            D.st[q] = 5 * L * pl->Ndense;
                        }
                        for (int q = 5; q < NF; q++) D.p[q] = nullptr, D.st[q] = 0;
                        Dp = &D;
                    }
                    ADS_TRY(pl->adj_level(A, pl->slot(w, n), pl->slot(w, n - 1), (int)n, pl->active[n],
                                          resid, Dp, w, sv, (int)pl->Pitch, 0, st));
                }
                return ADS_OK;
            }
        } ops{this, w, A, sv, resid, dense_hist, L, st};
        Reversal rv(w.M, w.S);
        ADS_TRY(rv.run(ops));
        dim3 g(cdiv(NY + 2, 256), NX + 2);
        el_grad_map<T><<<g, 256, 0, st>>>(w.grad, Npad, adj_groups() * Npad, (int)adj_groups(),
                                          (T *)grho, (T *)glam, (T *)gmu, (int)NX, (int)NY, (int)Pitch, sub);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int misfit(const WS &w, const T *traces, const T *obs, hipStream_t st)
    {
        launch_misfit<T>(traces, obs, w.resid, trace_elems, w.part, w.loss, st);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    // Full 5-field histories in the dense caller layout [shot][field][level][N]; the memory
    // variables live in a ping-pong scratch ([2][8][shot][N]).
    size_t history_scratch() const { return (size_t)(2 * 8 * nshots) * Ndense * sizeof(T) + 256; }
    int forward_history(const void *rho, const void *lam, const void *mu, const void *srcv,
                        void *hist, void *scratch, hipStream_t st)
    {
        const int64_t L = NMAX + 1, N = Ndense;
        ADS_HIP(hipMemsetAsync(hist, 0, (size_t)nshots * 5 * L * N * sizeof(T), st));
        T *mem = (T *)scratch;
        ADS_HIP(hipMemsetAsync(mem, 0, (size_t)2 * 8 * nshots * N * sizeof(T), st));
        auto lvl = [&](int64_t n) {
            El<T> e;
            for (int q = 0; q < 5; q++) {
                e.p[q] = (T *)hist + ((size_t)q * L + n) * N;
                e.st[q] = 5 * L * N;
            }
            for (int q = 0; q < 8; q++) {
                e.p[5 + q] = mem + ((size_t)(n & 1) * 8 + q) * nshots * N;
                e.st[5 + q] = N;
            }
            return e;
        };
        for (int64_t n = 1; n <= NMAX; n++)
            ADS_TRY(step(lvl(n - 1), lvl(n), (int)n, (int)nshots, (const T *)rho, (const T *)lam,
                         (const T *)mu, (const T *)srcv, nullptr, 0, (int)(NY + 2), 1, st));
        return ADS_OK;
    }
};

static int validate_params(const ads_elastic_params *p)
{
    if (!p) ADS_FAIL(ADS_ERR_ARG, "null params");
    if (p->NX < 3 || p->NY < 3 || p->NSTEP < 1)
        ADS_FAIL(ADS_ERR_SHAPE, "NX, NY must be >= 3 and NSTEP >= 1 (got %lld, %lld, %lld)",
                 (long long)p->NX, (long long)p->NY, (long long)p->NSTEP);
    if ((p->NX + 6) * (((p->NY + 6 + 63) / 64) * 64) >= (int64_t)1 << 31 || p->NX + 6 > 65535)
        ADS_FAIL(ADS_ERR_SHAPE, "grid too large");
    if (!(p->DELTAT > 0) || !(p->DELTAX > 0) || !(p->DELTAY > 0))
        ADS_FAIL(ADS_ERR_ARG, "DELTAT, DELTAX, DELTAY must be > 0");

    return ADS_OK;
}

template <typename T>
static int build_el_plan(const ads_elastic_params *p, const ads_shot_geometry *g, ElPlan<T> *P)
{
    ADS_TRY(validate_params(p));
    if (!g || g->nshots < 1 || !g->src_ptr || !g->rcv_ptr)
        ADS_FAIL(ADS_ERR_ARG, "geometry: need nshots >= 1, src_ptr and rcv_ptr");
    P->p = *p;
    // subdomain form: a block of NX x NY cells inside a 2-cell halo, stored as an
    // (NX+2) x (NY+2) "interior" with the usual 1-cell ring, i.e. (NX+4) x (NY+4) arrays
    const int sub = p->subdomain ? 1 : 0;
    P->sub = sub;
    P->PX = p->NX;
    P->PY = p->NY;
    P->NX = p->NX + 2 * sub;
    P->NY = p->NY + 2 * sub;
    P->Pitch = ((P->NY + 2 + 63) / 64) * 64;
    P->Npad = (P->NX + 2) * P->Pitch;
    P->Ndense = (P->NX + 2) * (P->NY + 2);
    P->nshots = g->nshots;
    P->dt_ = (T)p->DELTAT;
    P->ihx = (T)(1.0 / (24.0 * p->DELTAX));
    P->ihy = (T)(1.0 / (24.0 * p->DELTAY));
    const int64_t ns = g->nshots;
    if (g->src_ptr[0] != 0 || g->rcv_ptr[0] != 0) ADS_FAIL(ADS_ERR_ARG, "src_ptr/rcv_ptr must start at 0");
    std::vector<int64_t> nstep(ns);
    for (int64_t s = 0; s < ns; s++) {
        nstep[s] = g->nstep ? g->nstep[s] : p->NSTEP;
        if (nstep[s] < 1 || nstep[s] > p->NSTEP)
            ADS_FAIL(ADS_ERR_SHAPE, "shot %lld: nstep %lld outside [1, NSTEP]", (long long)s, (long long)nstep[s]);
        if (g->src_ptr[s + 1] < g->src_ptr[s] || g->rcv_ptr[s + 1] < g->rcv_ptr[s])
            ADS_FAIL(ADS_ERR_ARG, "src_ptr/rcv_ptr must be non-decreasing");
    }
    P->order.resize(ns);
    for (int64_t s = 0; s < ns; s++) P->order[s] = s;
    std::stable_sort(P->order.begin(), P->order.end(), [&](int64_t a, int64_t b) { return nstep[a] > nstep[b]; });
    P->NMAX = nstep[P->order[0]];
    P->nstep_sorted.resize(ns);
    for (int64_t q = 0; q < ns; q++) P->nstep_sorted[q] = nstep[P->order[q]];
    P->active.assign(P->NMAX + 2, 0);
    for (int64_t n = 0; n <= P->NMAX + 1; n++) {
        int c = 0;
        for (int64_t q = 0; q < ns; q++) c += P->nstep_sorted[q] >= n;
        P->active[n] = c;
    }
    P->srcv_off_caller.resize(ns);
    P->trace_off_caller.resize(ns);
    int64_t so = 0, to = 0;
    for (int64_t s = 0; s < ns; s++) {
        P->srcv_off_caller[s] = so;
        P->trace_off_caller[s] = to;
        so += nstep[s] * (g->src_ptr[s + 1] - g->src_ptr[s]);
        to += (nstep[s] + 1) * (g->rcv_ptr[s + 1] - g->rcv_ptr[s]);
    }
    P->srcv_elems = so;
    P->trace_elems = to;
    // 1-based index ranges: sources on the padded grid ([3, NX+2] x [3, NY+2], the block, in the
    // subdomain form: MPIElastic.jl:85-86 adds the halo offset 2); receivers on the padded grid,
    // or on the block itself in the subdomain form (MPIElastic.jl:110-111)
    const int64_t slo = 1 + 2 * sub, shi_x = P->NX + 2 - 2 * sub, shi_y = P->NY + 2 - 2 * sub;
    const int64_t rlo = 1, rhi_x = sub ? p->NX : P->NX + 2, rhi_y = sub ? p->NY : P->NY + 2;
    auto check = [&](const int64_t *I, const int64_t *J, const int64_t *Ty, int64_t e, const char *what) -> int {
        if (!Ty) ADS_FAIL(ADS_ERR_ARG, "%s types are null", what);
        int64_t i = I[e], j = J[e], t = Ty[e];
        const bool src = what[0] == 's';
        const int64_t lo = src ? slo : rlo, hx = src ? shi_x : rhi_x, hy = src ? shi_y : rhi_y;
        if (i < lo || i > hx || j < lo || j > hy)
            ADS_FAIL(ADS_ERR_INDEX, "%s %lld at (%lld, %lld) is outside [%lld, %lld] x [%lld, %lld]%s", what,
                     (long long)e, (long long)i, (long long)j, (long long)lo, (long long)hx, (long long)lo,
                     (long long)hy, sub ? " (subdomain block)" : " (padded grid)");
        if (t < 0 || t > 4)
            ADS_FAIL(ADS_ERR_ARG, "%s %lld has type %lld (must be 0..4: vx, vy, sxx, syy, sxy)", what,
                     (long long)e, (long long)t);
        return ADS_OK;
    };
    std::vector<int64_t> sp(ns + 1, 0), rp(ns + 1, 0), si, sj, st_, ri, rj, rt, nsrc(ns), soff(ns), toff(ns), nsv(ns);
    for (int64_t q = 0; q < ns; q++) {
        int64_t s = P->order[q];
        for (int64_t e = g->src_ptr[s]; e < g->src_ptr[s + 1]; e++) {
            if (!g->srci || !g->srcj) ADS_FAIL(ADS_ERR_ARG, "srci/srcj are null");
            ADS_TRY(check(g->srci, g->srcj, g->srctype, e, "source"));
            si.push_back(g->srci[e] - 1);
            sj.push_back(g->srcj[e] - 1);
            st_.push_back(g->srctype[e]);
        }
        for (int64_t e = g->rcv_ptr[s]; e < g->rcv_ptr[s + 1]; e++) {
            if (!g->rcvi || !g->rcvj) ADS_FAIL(ADS_ERR_ARG, "rcvi/rcvj are null");
            ADS_TRY(check(g->rcvi, g->rcvj, g->rcvtype, e, "receiver"));
            ri.push_back(g->rcvi[e] - 1 + 2 * sub);
            rj.push_back(g->rcvj[e] - 1 + 2 * sub);
            rt.push_back(g->rcvtype[e]);
        }
        sp[q + 1] = (int64_t)si.size();
        rp[q + 1] = (int64_t)ri.size();
        nsrc[q] = g->src_ptr[s + 1] - g->src_ptr[s];
        soff[q] = P->srcv_off_caller[s];
        toff[q] = P->trace_off_caller[s];
        nsv[q] = nstep[s];
    }
    const int64_t pitches[2] = {P->Pitch, P->NY + 2};
    for (int L = 0; L < 2; L++) {
        std::vector<int64_t> skey(si.size()), rkey(ri.size());
        for (size_t e = 0; e < si.size(); e++) skey[e] = (si[e] * pitches[L] + sj[e]) * 8 + st_[e];
        for (size_t e = 0; e < ri.size(); e++) rkey[e] = (ri[e] * pitches[L] + rj[e]) * 8 + rt[e];
        CellLists us = build_cell_lists(ns, sp.data(), skey), ur = build_cell_lists(ns, rp.data(), rkey);
        if (L == 0) {
            P->nkeys = us.shot_ptr[ns];
            ADS_TRY(P->d_us_ptr.upload(us.shot_ptr));
            ADS_TRY(P->d_us_eptr.upload(us.ent_ptr));
            ADS_TRY(P->d_us_ent.upload(us.ent));
            ADS_TRY(P->d_ur_ptr.upload(ur.shot_ptr));
            ADS_TRY(P->d_ur_eptr.upload(ur.ent_ptr));
            ADS_TRY(P->d_ur_ent.upload(ur.ent));
            P->max_ur = 0;
            for (int64_t q = 0; q < ns; q++) P->max_ur = std::max(P->max_ur, ur.shot_ptr[q + 1] - ur.shot_ptr[q]);
        }
        ADS_TRY(P->d_us_key[L].upload(us.cell));
        ADS_TRY(P->d_ur_key[L].upload(ur.cell));
        ADS_TRY(P->d_rcv_key[L].upload(rkey));
    }
    // CPML coefficients (Core.jl:231-407), host FP64 then cast
    std::vector<double> ax(2 * p->NX), bx(2 * p->NX), kx(2 * p->NX), ay(2 * p->NY), by(2 * p->NY), ky(2 * p->NY);
    ADS_TRY(ads_elastic_pml_params(p, ax.data(), bx.data(), kx.data(), ay.data(), by.data(), ky.data()));
    std::vector<T> prof;
    for (int64_t k = 0; k < 2 * p->NX; k++) prof.push_back((T)ax[k]);  // axf, axh
    for (int64_t k = 0; k < 2 * p->NX; k++) prof.push_back((T)bx[k]);  // bxf, bxh
    for (int64_t k = 0; k < 2 * p->NY; k++) prof.push_back((T)ay[k]);  // ayf, ayh
    for (int64_t k = 0; k < 2 * p->NY; k++) prof.push_back((T)by[k]);  // byf, byh
    for (int64_t r = -2; r <= P->NX + 3; r++) {  // packed per padded row
        const int64_t q = r - 1 - sub;
        const bool ok = r >= 1 + sub && r <= P->NX - sub;
        prof.push_back(ok ? (T)ax[q] : T(0));
        prof.push_back(ok ? (T)ax[p->NX + q] : T(0));
        prof.push_back(ok ? (T)bx[q] : T(0));
        prof.push_back(ok ? (T)bx[p->NX + q] : T(0));
    }
    for (int64_t j = 0; j <= P->NY + 1; j++) {  // packed per padded column
        const int64_t q = j - 1 - sub;
        const bool ok = j >= 1 + sub && j <= P->NY - sub;
        prof.push_back(ok ? (T)ay[q] : T(0));
        prof.push_back(ok ? (T)ay[p->NY + q] : T(0));
        prof.push_back(ok ? (T)by[q] : T(0));
        prof.push_back(ok ? (T)by[p->NY + q] : T(0));
    }
    ADS_TRY(P->d_prof.upload(prof));
    {  // the box where every a-profile vanishes: the memory variables stay 0 inside it
        std::vector<double> px((size_t)P->NX + 2, 1.0), py((size_t)P->NY + 2, 1.0);
        for (int64_t i = 1 + sub; i <= P->NX - sub; i++) {
            const int64_t q = i - 1 - sub;  // profile index
            px[i] = std::fabs(ax[q]) + std::fabs(ax[p->NX + q]);
        }
        for (int64_t j = 1 + sub; j <= P->NY - sub; j++) {
            const int64_t q = j - 1 - sub;
            py[j] = std::fabs(ay[q]) + std::fabs(ay[p->NY + q]);
        }
        P->mbox = quiet_box(px, py);
        P->nmv = P->mbox.cells();
    }
    int64_t cx = 0, cy = 0;
    for (int64_t k = 0; k < p->NX; k++) cx += (ax[k] != 0.0 || ax[p->NX + k] != 0.0);
    for (int64_t k = 0; k < p->NY; k++) cy += (ay[k] != 0.0 || ay[p->NY + k] != 0.0);
    P->f_x = (double)cx / (double)p->NX;
    P->f_y = (double)cy / (double)p->NY;
    {
        // plain rows/columns: every update range holds and the CPML is the identity there
        const int64_t NX = P->NX, NY = P->NY, po = 1 + sub;  // po: profile index offset
        std::vector<uint8_t> px(NX + 2, 0), py(NY + 2, 0);
        const T one = T(1), zero = T(0);
        for (int64_t i = 2; i <= NX - 1; i++)
            px[i] = (T)ax[i - po] == zero && (T)ax[p->NX + i - po] == zero && (T)bx[i - po] == one &&
                    (T)bx[p->NX + i - po] == one;
        for (int64_t j = 2; j <= NY - 1; j++)
            py[j] = (T)ay[j - po] == zero && (T)ay[p->NY + j - po] == zero && (T)by[j - po] == one &&
                    (T)by[p->NY + j - po] == one;
        for (int li = 0; li < 2; li++) {
            std::vector<int4> items;
            split_grid(px, py, NX, NY, li == 0 ? ElPlan<T>::VECP : 1, P->rect[li], items);
            P->nitems[li] = (int)items.size();
            ADS_TRY(P->d_items[li].upload(items));
            if (li == 0 && P->nitems[0] > 0) {
                std::vector<int4> sitems;
                frame_items(P->rect[0], NX, NY, sitems, FRS_ROWS, SW_FCOLS);
                P->nsitems = (int)sitems.size();
                ADS_TRY(P->d_sitems.upload(sitems));
            }
            ring_items(P->rect[li], 2, items);
            P->nring[li] = (int)items.size();
            if (!items.empty()) ADS_TRY(P->d_ring[li].upload(items));
        }
    }
    ADS_TRY(P->d_nsrc.upload(nsrc));
    ADS_TRY(P->d_srcv_off.upload(soff));
    ADS_TRY(P->d_rcv_ptr.upload(rp));
    ADS_TRY(P->d_trace_off.upload(toff));
    ADS_TRY(P->d_nstep.upload(nsv));
    return ADS_OK;
}

}  // namespace ads

using namespace ads;

struct ads_elastic_plan {
    std::unique_ptr<ElPlanBase> impl;
};

template <typename F> static int el_dispatch(ads_elastic_plan *pl, F &&f)
{
    if (!pl || !pl->impl) ADS_FAIL(ADS_ERR_ARG, "null plan");
    if (pl->impl->dtype == ADS_F64) return f(*static_cast<ElPlan<double> *>(pl->impl.get()));
    return f(*static_cast<ElPlan<float> *>(pl->impl.get()));
}

extern "C" {

int ads_elastic_plan_create(const ads_elastic_params *p, const ads_shot_geometry *g, int dtype,
                            ads_elastic_plan **out)
{
    clear_error();
    if (!out) ADS_FAIL(ADS_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (dtype != ADS_F32 && dtype != ADS_F64) ADS_FAIL(ADS_ERR_ARG, "dtype must be ADS_F32 or ADS_F64");
    auto pl = std::make_unique<ads_elastic_plan>();
    int rc;
    if (dtype == ADS_F64) {
        auto P = std::make_unique<ElPlan<double>>();
        P->dtype = dtype;
        rc = build_el_plan<double>(p, g, P.get());
        pl->impl = std::move(P);
    } else {
        auto P = std::make_unique<ElPlan<float>>();
        P->dtype = dtype;
        rc = build_el_plan<float>(p, g, P.get());
        pl->impl = std::move(P);
    }
    if (rc != ADS_OK) return rc;
    *out = pl.release();
    return ADS_OK;
}

int ads_elastic_plan_destroy(ads_elastic_plan *plan)
{
    delete plan;
    return ADS_OK;
}

int ads_elastic_plan_sizes(const ads_elastic_plan *plan, int64_t *srcv_elems, int64_t *trace_elems,
                           int64_t *srcv_off, int64_t *trace_off)
{
    clear_error();
    if (!plan || !plan->impl) ADS_FAIL(ADS_ERR_ARG, "null plan");
    const ElPlanBase &b = *plan->impl;
    if (srcv_elems) *srcv_elems = b.srcv_elems;
    if (trace_elems) *trace_elems = b.trace_elems;
    for (int64_t s = 0; s < b.nshots; s++) {
        if (srcv_off) srcv_off[s] = b.srcv_off_caller[s];
        if (trace_off) trace_off[s] = b.trace_off_caller[s];
    }
    return ADS_OK;
}

size_t ads_elastic_plan_workspace_size(const ads_elastic_plan *plan, int64_t K)
{
    if (!plan || !plan->impl) return 0;
    if (plan->impl->dtype == ADS_F64) return static_cast<ElPlan<double> *>(plan->impl.get())->ws_size(K);
    return static_cast<ElPlan<float> *>(plan->impl.get())->ws_size(K);
}

int64_t ads_elastic_plan_choose_interval(const ads_elastic_plan *plan, size_t budget)
{
    if (!plan || !plan->impl) return 0;
    // as for the acoustic plan: fewest segments, then the shortest interval with that count
    const int64_t L = std::max<int64_t>(plan->impl->NMAX, 1);
    for (int64_t M = 1; M <= L; M++) {
        const int64_t K = (L + M - 1) / M;
        const size_t need = plan->impl->dtype == ADS_F64
                                ? static_cast<ElPlan<double> *>(plan->impl.get())->ws_size(K, -1)
                                : static_cast<ElPlan<float> *>(plan->impl.get())->ws_size(K, -1);
        if (need <= budget) return K;
        if (K == 1) break;
    }
    return 0;
}

int ads_elastic_plan_set_snapshots(ads_elastic_plan *plan, int64_t snapshots)
{
    clear_error();
    if (!plan || !plan->impl) ADS_FAIL(ADS_ERR_ARG, "null plan");
    plan->impl->snapshots = snapshots < 0 ? -1 : snapshots;
    return ADS_OK;
}

int64_t ads_elastic_plan_snapshots(const ads_elastic_plan *plan)
{
    return plan && plan->impl ? plan->impl->snapshots : -1;
}

int ads_elastic_plan_choose_schedule(ads_elastic_plan *plan, size_t budget, int64_t *K, int64_t *S)
{
    clear_error();
    if (!plan || !plan->impl || !K || !S) ADS_FAIL(ADS_ERR_ARG, "null argument");
    const int rc = plan->impl->dtype == ADS_F64
                       ? static_cast<ElPlan<double> *>(plan->impl.get())->choose_schedule(budget, K, S)
                       : static_cast<ElPlan<float> *>(plan->impl.get())->choose_schedule(budget, K, S);
    if (rc != ADS_OK) ADS_FAIL(rc, "no checkpoint schedule fits %zu bytes", budget);
    plan->impl->snapshots = *S;
    return ADS_OK;
}

int ads_elastic_plan_set_adjoint_kernels(ads_elastic_plan *plan, int streaming, int rows, int shots)
{
    clear_error();
    if (!plan || !plan->impl) ADS_FAIL(ADS_ERR_ARG, "null plan");
    if (streaming < 0 || streaming > 2 || rows < 1 || rows > SW_CH_MAX || shots < 1 || shots > SW_NW ||
        rows * shots > SW_LDS_ROWS)
        ADS_FAIL(ADS_ERR_ARG, "streaming must be 0, 1 or 2, rows in [1, %d], shots in [1, %d] and "
                 "rows x shots <= %d", SW_CH_MAX, SW_NW, SW_LDS_ROWS);
    plan->impl->strm = streaming;
    plan->impl->strm_ch = rows;
    plan->impl->asg = shots;
    return ADS_OK;
}

int ads_elastic_plan_set_forward_kernels(ads_elastic_plan *plan, int fused, int rows)
{
    clear_error();
    if (!plan || !plan->impl) ADS_FAIL(ADS_ERR_ARG, "null plan");
    if (fused < 0 || fused > 1 || rows < 0 || rows > 1024)
        ADS_FAIL(ADS_ERR_ARG, "fused must be 0 or 1 and rows in [0, 1024]");
    plan->impl->ffused = fused;
    plan->impl->fch = rows;
    return ADS_OK;
}

int ads_elastic_plan_forward(ads_elastic_plan *plan, const void *rho, const void *lam,
                             const void *mu, const void *srcv, void *traces, void *ws, size_t wsb,
                             int64_t K, ads_stream_t stream)
{
    clear_error();
    if (!rho || !lam || !mu || !srcv) ADS_FAIL(ADS_ERR_ARG, "null argument");
    return el_dispatch(plan, [&](auto &P) { return P.forward(rho, lam, mu, srcv, traces, ws, wsb, K, S(stream)); });
}

int ads_elastic_plan_backward(ads_elastic_plan *plan, const void *rho, const void *lam,
                              const void *mu, const void *srcv, const void *grad_traces,
                              void *grad_rho, void *grad_lam, void *grad_mu, void *ws, size_t wsb,
                              int64_t K, ads_stream_t stream)
{
    clear_error();
    if (!rho || !lam || !mu || !srcv || !grad_traces || !grad_rho || !grad_lam || !grad_mu)
        ADS_FAIL(ADS_ERR_ARG, "null argument");
    return el_dispatch(plan, [&](auto &P) {
        using R = typename std::remove_const<decltype(P.dt_)>::type;
        return P.backward(srcv, (const R *)grad_traces, nullptr, grad_rho, grad_lam, grad_mu, ws, wsb, K, S(stream));
    });
}

int ads_elastic_plan_gradient(ads_elastic_plan *plan, const void *rho, const void *lam,
                              const void *mu, const void *srcv, const void *observed, double *loss,
                              void *grad_rho, void *grad_lam, void *grad_mu, void *traces, void *ws,
                              size_t wsb, int64_t K, ads_stream_t stream)
{
    clear_error();
    if (!rho || !lam || !mu || !srcv || !observed || !grad_rho || !grad_lam || !grad_mu)
        ADS_FAIL(ADS_ERR_ARG, "null argument");
    return el_dispatch(plan, [&](auto &P) -> int {
        using R = typename std::remove_const<decltype(P.dt_)>::type;
        typename std::remove_reference<decltype(P)>::type::WS w;
        ADS_TRY(P.check_ws(w, ws, wsb, K));
        R *tr = traces ? (R *)traces : w.resid;
        ADS_TRY(P.forward(rho, lam, mu, srcv, tr, ws, wsb, K, S(stream)));
        ADS_TRY(P.misfit(w, tr, (const R *)observed, S(stream)));
        if (loss) ADS_HIP(hipMemcpyAsync(loss, w.loss, sizeof(double), hipMemcpyDeviceToDevice, S(stream)));
        return P.backward(srcv, w.resid, nullptr, grad_rho, grad_lam, grad_mu, ws, wsb, K, S(stream));
    });
}

int ads_elastic_plan_backward_dense(ads_elastic_plan *plan, const void *rho, const void *lam,
                                    const void *mu, const void *srcv, const void *grad_hist,
                                    void *grad_rho, void *grad_lam, void *grad_mu, void *ws,
                                    size_t wsb, int64_t K, ads_stream_t stream)
{
    clear_error();
    if (!rho || !lam || !mu || !srcv || !grad_hist || !grad_rho || !grad_lam || !grad_mu)
        ADS_FAIL(ADS_ERR_ARG, "null argument");
    return el_dispatch(plan, [&](auto &P) -> int {
        if (P.NMAX != P.p.NSTEP) ADS_FAIL(ADS_ERR_SHAPE, "dense cotangents need nstep == NSTEP for every shot");
        ADS_TRY(P.forward(rho, lam, mu, srcv, nullptr, ws, wsb, K, S(stream)));
        return P.backward(srcv, nullptr, grad_hist, grad_rho, grad_lam, grad_mu, ws, wsb, K, S(stream));
    });
}

size_t ads_elastic_plan_history_scratch_size(const ads_elastic_plan *plan)
{
    if (!plan || !plan->impl) return 0;
    if (plan->impl->dtype == ADS_F64) return static_cast<ElPlan<double> *>(plan->impl.get())->history_scratch();
    return static_cast<ElPlan<float> *>(plan->impl.get())->history_scratch();
}

int ads_elastic_forward_history(ads_elastic_plan *plan, const void *rho, const void *lam,
                                const void *mu, const void *srcv, void *hist, void *scratch,
                                ads_stream_t stream)
{
    clear_error();
    if (!rho || !lam || !mu || !srcv || !hist || !scratch) ADS_FAIL(ADS_ERR_ARG, "null argument");
    return el_dispatch(plan, [&](auto &P) -> int {
        if (P.NMAX != P.p.NSTEP) ADS_FAIL(ADS_ERR_SHAPE, "history mode needs nstep == NSTEP for every shot");
        return P.forward_history(rho, lam, mu, srcv, hist, scratch, S(stream));
    });
}

}  // extern "C"
