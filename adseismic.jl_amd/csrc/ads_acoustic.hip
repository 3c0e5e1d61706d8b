// ads_acoustic.hip -- acoustic PML time stepper for gfx950: forward, hand-written adjoint,
// checkpointed multi-shot batches, full-history drop-in, kernel-shape autotuner.
//
// Scheme (src/Core.jl:528-549 / AcousticOneStepCpu.h:1-48), interior cell (i, j):
//   u^n = [(2 - s t dt^2) w + C lap(w) + ex (phi_{i+1} - phi_{i-1}) + ey (psi_{j+1} - psi_{j-1})
//          - (1 - (s+t) dt/2) wold] / (1 + (s+t) dt/2),
//   lap(w) = cx (w_{i+1} + w_{i-1} - 2w) + cy (w_{j+1} + w_{j-1} - 2w),  cx = dt^2/hx^2
//   phi^n = (1 - dt s) phi + C (t - s) fx (v_{i+1} - v_{i-1}),  fx = dt/(2hx)
//   psi^n = (1 - dt t) psi + C (s - t) fy (v_{j+1} - v_{j-1})
// with w = u^{n-1}, wold = u^{n-2}, v = u^n before source injection (PropagatorKernel 0) or
// v = w (PropagatorKernel 1/2).  s = sigma_x(i), t = tau_y(j) (separable, Core.jl:738-761).
// phi/psi are non-zero only where s or t is (the PML strips); interior cells skip them.
//
// Adjoint (exact transpose of the discrete loop; P^n = dL/du^n / d, d = 1 + (s+t) dt/2):
//   dL/du^n = R^n + (2 - s t dt^2) P^{n+1} + lap(C P^{n+1}) - (1 - (s+t) dt/2) P^{n+2}
//             + fx [Cp_{i-1} F_{i-1} - Cp_{i+1} F_{i+1}] + fy [Cq_{j-1} G_{j-1} - Cq_{j+1} G_{j+1}]
//   F^n = (1 - dt s) F^{n+1} + ex (P^{n+1}_{i-1} - P^{n+1}_{i+1})    (dL/dphi^n), G likewise
//   dL/dC += P^n lap(u^{n-1}) + F^n (t - s) fx (v_{i+1} - v_{i-1}) + G^n (s - t) fy (...)
// where Cp = C (t - s), Cq = C (s - t) and F in the dL/du^n line is F^{n+1} (kernel 1) or F^n
// (kernel 0).  Receiver residuals R^n are added after the stencil pass by a per-cell gather.
//
// Kernel shape.  A workgroup is 4 wavefronts; each lane owns VEC consecutive depth cells (j)
// (one 16-byte vector per row in the 64-element-pitched plan layout) of RPT consecutive rows
// (i).  j-1/j+1 come from neighbouring lanes by shuffle; the two edge lanes load theirs.  A
// workgroup loops over a group of Sg shots, SB at a time, issuing every load of those SB shots
// before the first store; the model rows stay in registers across the group and, in the
// adjoint, dL/dC is accumulated in registers and written once per group.  PML work is
// lane-masked: only lanes with a cell near the PML read phi/psi (F/G), with vector loads issued
// together with the others, and only wavefronts with such a lane evaluate the PML formula.
// Workgroup ids are swizzled so that each of the 8 XCDs walks a contiguous band of tiles.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "ads_common.h"
#include "ads_schedule.h"
#include "ads_stencil.h"

namespace ads {

template <typename T> struct AcCoef {
    T cx, cy, ex, ey, fx, fy, dt, hdt, dt2;
};

template <typename T> AcCoef<T> make_coef(double dt, double hx, double hy)
{
    AcCoef<T> k;
    k.cx = (T)(dt * dt / (hx * hx));
    k.cy = (T)(dt * dt / (hy * hy));
    k.ex = (T)(dt * dt / (2.0 * hx));
    k.ey = (T)(dt * dt / (2.0 * hy));
    k.fx = (T)(dt / (2.0 * hx));
    k.fy = (T)(dt / (2.0 * hy));
    k.dt = (T)dt;
    k.hdt = (T)(dt / 2.0);
    k.dt2 = (T)(dt * dt);
    return k;
}

// Field addressing: level pointer + shot * stride.
template <typename T> struct Lv {
    T *p;
    int64_t stride;
    __device__ __host__ T *at(int s) const { return p + (int64_t)s * stride; }
};
template <typename T> inline Lv<const T> cst(Lv<T> a) { return Lv<const T>{a.p, a.stride}; }

template <typename T> __device__ inline T ldg(const T *p, int64_t i, bool ok) { return ok ? p[i] : T(0); }

// ---------------------------------------------------------------------------------------------
// forward step: u^n (pre-source) and, for kernel 1, phi^n/psi^n
// ---------------------------------------------------------------------------------------------
// Sources added by the frame work of a level (PropagatorKernel 1, no source in the interior
// rectangle): per cell, cell_idx is -1 or a source cell whose entries [ptr[c], ptr[c+1]) are
// (shot, srcv offset of the entry's column, sources of that shot) in (shot, source) order.
template <typename T> struct SrcIn {
    const int32_t *cell_idx, *ptr, *shot;
    const int64_t *base, *ns;
    const T *srcv;
    int n;  // level
};

template <typename T> struct FwdArgs {
    Lv<T> u;
    Lv<const T> w, wold, phi, psi;
    Lv<T> phio, psio;
    const T *C, *sx, *ty;
    const uint8_t *nearx, *neary;
    int NX, NY, P, act, Sg;
    TileMap map;
    AcCoef<T> k;
    SrcIn<T> src;  // cell_idx null: sources are added by ac_src_rcv
};

// u^n + dt^2 srcv[n-1, e] over the cell's entries of shot s, in source order (Core.jl:593-601)
template <typename T>
__device__ inline T add_sources(const SrcIn<T> &q, int c, int s, T v, T dt2)
{
    for (int e = q.ptr[c]; e < q.ptr[c + 1]; e++)
        if (q.shot[e] == s) v += q.srcv[q.base[e] + (int64_t)(q.n - 1) * q.ns[e]] * dt2;
    return v;
}

// Rectangle forward: u^n = 2 w - wold + C lap(w).  A wavefront owns RPT rows x (64 lanes x VEC
// depth cells); a workgroup loops over a group of Sg shots, SB at a time, issuing every load of
// those SB shots before the first store, so each wavefront keeps SB shots' rows in flight.
template <typename T, int RPT, int VEC, int SB>
__device__ inline void fwd_rect_body(const FwdArgs<T> &a, const Rect &R, int bid)
{
    int g, strip, tj;
    if (!decode_tile(a.map, g, strip, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int j0 = R.jL + (tj * WAVE + lane) * VEC;
    const int ib = R.iL + (strip * NWAVES + wv) * RPT;
    if (ib >= R.iR) return;  // wave-uniform
    const int P = a.P, nr = min(RPT, R.iR - ib);
    const bool act = j0 < R.jR;
    const bool le = lane == 0 && act;                                // reads j0-1 >= jL-1 >= 0
    const bool re = act && (lane == WAVE - 1 || j0 + VEC == R.jR);  // reads j0+VEC <= jR <= NY+1
    const AcCoef<T> k = a.k;
    T Cr[RPT][VEC];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        if (act && r < nr) vld(Cr[r], a.C + (int64_t)(ib + r) * P + j0);
        else vzero(Cr[r]);
    }
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int s = g * a.Sg; s < s1; s += SB) {
        T W[SB][RPT + 2][VEC], WO[SB][RPT][VEC], EL[SB][RPT], ER[SB][RPT];
#pragma unroll
        for (int b = 0; b < SB; b++) {
            if (s + b >= s1) break;  // uniform
            const T *__restrict__ w = a.w.at(s + b);
            const T *__restrict__ wo = a.wold.at(s + b);
#pragma unroll
            for (int q = 0; q < RPT + 2; q++) {  // rows iL-1 .. iR lie inside [0, NX+1]
                if (act && q <= nr + 1) vld(W[b][q], w + (int64_t)(ib - 1 + q) * P + j0);
                else vzero(W[b][q]);
            }
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                const int64_t row = (int64_t)(ib + r) * P;
                if (act && r < nr) vld(WO[b][r], wo + row + j0);
                else vzero(WO[b][r]);
                EL[b][r] = le && r < nr ? w[row + j0 - 1] : T(0);
                ER[b][r] = re && r < nr ? w[row + j0 + VEC] : T(0);
            }
        }
#pragma unroll
        for (int b = 0; b < SB; b++) {
            if (s + b >= s1) break;
            T *__restrict__ u = a.u.at(s + b);
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                if (r >= nr) break;  // uniform
                T wl, wr;
                lr_of(W[b][r + 1], EL[b][r], ER[b][r], lane, re, wl, wr);
                T out[VEC];
#pragma unroll
                for (int e = 0; e < VEC; e++) {
                    const T wc = W[b][r + 1][e];
                    const T ww = e == 0 ? wl : W[b][r + 1][e - 1];
                    const T we = e == VEC - 1 ? wr : W[b][r + 1][e + 1];
                    const T lap = k.cx * (W[b][r + 2][e] + W[b][r][e] - T(2) * wc) + k.cy * (we + ww - T(2) * wc);
                    out[e] = T(2) * wc - WO[b][r][e] + Cr[r][e] * lap;
                }
                if (act) vst(u + (int64_t)(ib + r) * P + j0, out);
            }
        }
    }
}

// Two-level rectangle forward (temporal blocking).  From w = u^{n-1} and wold = u^{n-2} one pass
// writes u^n on the rectangle R and u^{n+1} on R shrunk by one cell; the ring of R at level n+1
// is left to the frame work (the two-level frame body in the same launch, or a frame launch for
// level n+1 with kernel 0).  A wavefront owns RPT rows of u^{n+1}: it recomputes u^n
// on one halo row above and below and, on its segment's edge lanes, one halo column on each
// side, so no value crosses wavefronts.  Every value is the single-level formula in the same
// order, so results equal two single-level steps bit for bit.  HBM per cell and two levels:
// read u^{n-1} and u^{n-2}, write u^n and u^{n+1} (two thirds of the single-level traffic).
// Valid only when no source lies in R: level n must carry its sources before level n+1 reads it,
// and the sources are added by the frame work or between the launches of a pair.
template <typename T, int RPT, int VEC>
__device__ inline void fwd_rect2_body(const FwdArgs<T> &a, const Lv<T> &u2, const Rect &R, int bid)
{
    int g, strip, tj;
    if (!decode_tile(a.map, g, strip, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int ja = R.jL + tj * WAVE * VEC;  // first column of the segment
    const int j0 = ja + lane * VEC;
    const int ib = R.iL + (strip * NWAVES + wv) * RPT;
    if (ib >= R.iR) return;  // wave-uniform
    const int P = a.P, nr = min(RPT, R.iR - ib);
    const bool act = j0 < R.jR;
    const int jb = min(ja + WAVE * VEC, R.jR);  // one past the segment's last column
    const bool le = lane == 0;                    // ja < R.jR, so lane 0 is active
    const bool re = act && (lane == WAVE - 1 || j0 + VEC == R.jR);
    const bool hl = le && ja - 1 >= R.jL;  // the halo column left of the segment lies in R
    const bool hr = re && jb < R.jR;       // the halo column right of it
    const AcCoef<T> k = a.k;
    constexpr int NW = RPT + 4, NU = RPT + 2;  // w rows ib-2+q; u^n, wold, C rows ib-1+q
    bool urow[NU];
    T Cr[NU][VEC], CL[NU], CR[NU];
#pragma unroll
    for (int q = 0; q < NU; q++) {
        const int i = ib - 1 + q;
        urow[q] = q <= nr + 1 && i >= R.iL && i < R.iR;
        const int64_t o = (int64_t)i * P;
        if (act && urow[q]) vld(Cr[q], a.C + o + j0);
        else vzero(Cr[q]);
        CL[q] = hl && urow[q] ? a.C[o + ja - 1] : T(0);
        CR[q] = hr && urow[q] ? a.C[o + jb] : T(0);
    }
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int s = g * a.Sg; s < s1; s++) {
        const T *__restrict__ w = a.w.at(s);
        const T *__restrict__ wo = a.wold.at(s);
        T W[NW][VEC], WL1[NW], WR1[NW];  // w rows; w at columns ja-1 and jb
        T WO[NU][VEC], WL2[NU], WR2[NU], WOL[NU], WOR[NU];
#pragma unroll
        for (int q = 0; q < NW; q++) {  // rows R.iL-1 .. R.iR lie inside [0, NX+1]
            const int x = ib - 2 + q;
            const bool ok = q <= nr + 3 && x >= R.iL - 1 && x <= R.iR;
            const int64_t o = (int64_t)x * P;
            if (act && ok) vld(W[q], w + o + j0);
            else vzero(W[q]);
            WL1[q] = le && ok ? w[o + ja - 1] : T(0);
            WR1[q] = re && ok ? w[o + jb] : T(0);
        }
#pragma unroll
        for (int q = 0; q < NU; q++) {
            const int64_t o = (int64_t)(ib - 1 + q) * P;
            if (act && urow[q]) vld(WO[q], wo + o + j0);
            else vzero(WO[q]);
            WL2[q] = hl && urow[q] ? w[o + ja - 2] : T(0);
            WR2[q] = hr && urow[q] ? w[o + jb + 1] : T(0);
            WOL[q] = hl && urow[q] ? wo[o + ja - 1] : T(0);
            WOR[q] = hr && urow[q] ? wo[o + jb] : T(0);
        }
        // u^n on rows ib-1 .. ib+nr (and the halo columns on the edge lanes)
        T U[NU][VEC], UL[NU], UR[NU];
#pragma unroll
        for (int q = 0; q < NU; q++) {
            T wl, wr;
            lr_of(W[q + 1], WL1[q + 1], WR1[q + 1], lane, re, wl, wr);
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const T wc = W[q + 1][e];
                const T ww = e == 0 ? wl : W[q + 1][e - 1];
                const T we = e == VEC - 1 ? wr : W[q + 1][e + 1];
                const T lap = k.cx * (W[q + 2][e] + W[q][e] - T(2) * wc) + k.cy * (we + ww - T(2) * wc);
                U[q][e] = T(2) * wc - WO[q][e] + Cr[q][e] * lap;
            }
            {
                const T wc = WL1[q + 1];
                const T lap = k.cx * (WL1[q + 2] + WL1[q] - T(2) * wc) + k.cy * (W[q + 1][0] + WL2[q] - T(2) * wc);
                UL[q] = T(2) * wc - WOL[q] + CL[q] * lap;
            }
            {
                const T wc = WR1[q + 1];
                const T lap = k.cx * (WR1[q + 2] + WR1[q] - T(2) * wc) + k.cy * (WR2[q] + W[q + 1][VEC - 1] - T(2) * wc);
                UR[q] = T(2) * wc - WOR[q] + CR[q] * lap;
            }
        }
        T *__restrict__ u1 = a.u.at(s);
        T *__restrict__ up = u2.at(s);
#pragma unroll
        for (int q = 1; q <= RPT; q++) {
            if (q > nr) break;  // uniform
            const int i = ib - 1 + q;
            const int64_t o = (int64_t)i * P + j0;
            if (act) vst(u1 + o, U[q]);
            if (i <= R.iL || i >= R.iR - 1) continue;  // uniform: u^{n+1} only inside R shrunk by 1
            T ul, ur;
            lr_of(U[q], UL[q], UR[q], lane, re, ul, ur);
            T out[VEC];
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const T uc = U[q][e];
                const T uw = e == 0 ? ul : U[q][e - 1];
                const T ue = e == VEC - 1 ? ur : U[q][e + 1];
                const T lap = k.cx * (U[q + 1][e] + U[q - 1][e] - T(2) * uc) + k.cy * (ue + uw - T(2) * uc);
                out[e] = T(2) * uc - W[q + 1][e] + Cr[q][e] * lap;
            }
            if (!act) continue;
            if (j0 > R.jL && j0 + VEC < R.jR) {
                vst(up + o, out);
            } else {
#pragma unroll
                for (int e = 0; e < VEC; e++)
                    if (j0 + e > R.jL && j0 + e < R.jR - 1) up[o + e] = out[e];
            }
        }
    }
}

// Two-level rectangle forward, streaming form: a wavefront owns ROWS rows of one column
// segment for one shot (a workgroup takes NWAVES shots) and walks down them, holding three rows
// of u^{n-1}, u^n and C in registers and loading the next rows one step ahead, so each row of
// u^{n-1} and u^{n-2} is read about once instead of RPT+4 and RPT+2 times per RPT rows.  The
// values and writes are those of fwd_rect2_body (bit-identical).
template <typename T, int VEC>
struct Rect2Row {  // one row of the window: the vector and the segment's halo columns
    T v[VEC];
    T l, r;
};
template <typename T, int VEC, int ROWS>
__device__ inline void fwd_rect2s_body(const FwdArgs<T> &a, const Lv<T> &u2, const Rect &R, int bid)
{
    int g, strip, tj;
    if (!decode_tile(a.map, g, strip, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int s = g * NWAVES + wv;
    if (s >= a.act) return;  // wave-uniform; no barrier below
    const int ja = R.jL + tj * WAVE * VEC;  // first column of the segment
    const int j0 = ja + lane * VEC;
    const int ib = R.iL + strip * ROWS;  // first owned row
    const int ie = min(ib + ROWS, R.iR);  // one past the last owned row
    const int P = a.P;
    const bool act = j0 < R.jR;
    const int jb = min(ja + WAVE * VEC, R.jR);
    const bool le = lane == 0;
    const bool re = act && (lane == WAVE - 1 || j0 + VEC == R.jR);
    const bool hl = le && ja - 1 >= R.jL;
    const bool hr = re && jb < R.jR;
    const AcCoef<T> k = a.k;
    const T *__restrict__ w = a.w.at(s);
    const T *__restrict__ wo = a.wold.at(s);
    T *__restrict__ u1 = a.u.at(s);
    T *__restrict__ up = u2.at(s);
    // byte offsets: row x starts at x * pb, this lane's vector at jo
    const uint32_t pb = (uint32_t)P * (uint32_t)sizeof(T), jo = (uint32_t)j0 * (uint32_t)sizeof(T);
    // u^{n-1} row x (rows R.iL-1 .. R.iR lie inside [0, NX+1]) with columns ja-1 and jb
    auto loadW = [&](int x, Rect2Row<T, VEC> &o) {
        const bool ok = x >= R.iL - 1 && x <= R.iR;
        const uint32_t q = (uint32_t)x * pb;
        if (act && ok) vld(o.v, at_off(w, q + jo));
        else vzero(o.v);
        o.l = le && ok ? at_off(w, q)[ja - 1] : T(0);
        o.r = re && ok ? at_off(w, q)[jb] : T(0);
    };
    // inputs of level n on row x besides u^{n-1}: u^{n-2}, C, and for the halo columns
    // u^{n-1} at ja-2 / jb+1, u^{n-2} and C at ja-1 / jb
    struct In {
        T wo[VEC], c[VEC];
        T cl, cr, wl2, wr2, wol, wor;
    };
    auto loadIn = [&](int x, In &o) {
        const bool ok = x >= R.iL && x < R.iR;
        const uint32_t q = (uint32_t)x * pb;
        if (act && ok) {
            vld(o.wo, at_off(wo, q + jo));
            vld(o.c, at_off(a.C, q + jo));
        } else {
            vzero(o.wo);
            vzero(o.c);
        }
        o.cl = hl && ok ? at_off(a.C, q)[ja - 1] : T(0);
        o.cr = hr && ok ? at_off(a.C, q)[jb] : T(0);
        o.wl2 = hl && ok ? at_off(w, q)[ja - 2] : T(0);
        o.wr2 = hr && ok ? at_off(w, q)[jb + 1] : T(0);
        o.wol = hl && ok ? at_off(wo, q)[ja - 1] : T(0);
        o.wor = hr && ok ? at_off(wo, q)[jb] : T(0);
    };
    // level n on row x from u^{n-1} rows x-1, x, x+1
    auto level_n = [&](const Rect2Row<T, VEC> &Wn, const Rect2Row<T, VEC> &Wc,
                       const Rect2Row<T, VEC> &Ws, const In &I, Rect2Row<T, VEC> &U) {
        T wl, wr;
        lr_of(Wc.v, Wc.l, Wc.r, lane, re, wl, wr);
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            const T wc = Wc.v[e];
            const T ww = e == 0 ? wl : Wc.v[e - 1];
            const T we = e == VEC - 1 ? wr : Wc.v[e + 1];
            const T lap = k.cx * (Ws.v[e] + Wn.v[e] - T(2) * wc) + k.cy * (we + ww - T(2) * wc);
            U.v[e] = T(2) * wc - I.wo[e] + I.c[e] * lap;
        }
        {
            const T wc = Wc.l;
            const T lap = k.cx * (Ws.l + Wn.l - T(2) * wc) + k.cy * (Wc.v[0] + I.wl2 - T(2) * wc);
            U.l = T(2) * wc - I.wol + I.cl * lap;
        }
        {
            const T wc = Wc.r;
            const T lap = k.cx * (Ws.r + Wn.r - T(2) * wc) + k.cy * (I.wr2 + Wc.v[VEC - 1] - T(2) * wc);
            U.r = T(2) * wc - I.wor + I.cr * lap;
        }
    };
    // window: W0 = row x-1, W1 = row x, W2 = row x+1 (u^{n-1}); Wn = row x+2 (next load);
    // U0 = level n on row x-2, U1 on row x-1; C1 = C on row x-1; I1 = inputs of row x, In_ next
    Rect2Row<T, VEC> W0, W1, W2, U0, U1;
    T C1[VEC];
    In I1;
    const int x0 = ib - 1;  // first level-n row
    loadW(x0 - 1, W0);
    loadW(x0, W1);
    loadW(x0 + 1, W2);
    loadIn(x0, I1);
    vzero(U0.v);
    vzero(U1.v);
    U0.l = U0.r = U1.l = U1.r = T(0);
    vzero(C1);
#pragma unroll 1
    for (int x = x0; x <= ie; x++) {
        // next rows first, so their latency overlaps this step
        Rect2Row<T, VEC> W3;
        In I2;
        loadW(x + 2, W3);
        loadIn(x + 1, I2);
        Rect2Row<T, VEC> U2;  // level n on row x
        level_n(W0, W1, W2, I1, U2);
        if (x >= ib && x < ie && act) vst(at_off(u1, (uint32_t)x * pb + jo), U2.v);
        const int y = x - 1;  // level n+1 on row y from level n rows y-1, y, y+1
        if (y >= ib && y > R.iL && y < R.iR - 1) {
            T ul, ur;
            lr_of(U1.v, U1.l, U1.r, lane, re, ul, ur);
            T out[VEC];
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const T uc = U1.v[e];
                const T uw = e == 0 ? ul : U1.v[e - 1];
                const T ue = e == VEC - 1 ? ur : U1.v[e + 1];
                const T lap = k.cx * (U2.v[e] + U0.v[e] - T(2) * uc) + k.cy * (ue + uw - T(2) * uc);
                out[e] = T(2) * uc - W0.v[e] + C1[e] * lap;
            }
            if (act) {
                T *__restrict__ po = at_off(up, (uint32_t)y * pb + jo);
                if (j0 > R.jL && j0 + VEC < R.jR) {
                    vst(po, out);
                } else {
#pragma unroll
                    for (int e = 0; e < VEC; e++)
                        if (j0 + e > R.jL && j0 + e < R.jR - 1) po[e] = out[e];
                }
            }
        }
        // rotate
        W0 = W1;
        W1 = W2;
        W2 = W3;
        U0 = U1;
        U1 = U2;
#pragma unroll
        for (int e = 0; e < VEC; e++) C1[e] = I1.c[e];
        I1 = I2;
    }
}

// Frame forward, one depth cell per lane.  A workgroup takes one frame item and its four
// wavefronts take every fourth shot of the group, so the thin PML strips still fill the GPU.
// Only lanes with a cell near the PML read phi/psi; cells of the padding ring get u = 0
// (sources are added afterwards).
template <typename T, int V>
__device__ inline void fwd_frame_body(const FwdArgs<T> &a, const TileMap &map,
                                      const int4 *__restrict__ items, int nitems, int bid)
{
    int g, blk, tj;
    if (!decode_tile(map, g, blk, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int it = blk;  // one item per workgroup; its 4 wavefronts take every 4th shot
    if (it >= nitems) return;  // workgroup-uniform
    const int4 m = items[it];
    const int ib = m.x, nr = m.y, j = m.z + lane;
    const int NX = a.NX, NY = a.NY, P = a.P;
    const bool act = j < m.w;
    const bool le = lane == 0 && act && j >= 1;
    const bool re = act && (lane == WAVE - 1 || j + 1 == m.w);
    const bool reok = re && j + 1 <= NY + 1;
    const AcCoef<T> k = a.k;
    const bool jin = j >= 1 && j <= NY;
    const T tu = act ? a.ty[j] : T(0);
    const bool ny = act && a.neary[j] != 0;
    bool rin[FRPT], nx[FRPT];
    T sg[FRPT], Cr[FRPT];
    bool lp = false;
#pragma unroll
    for (int r = 0; r < FRPT; r++) {
        const int i = ib + r;
        const bool ok = r < nr;
        rin[r] = ok && i >= 1 && i <= NX;
        nx[r] = ok && a.nearx[i] != 0;
        sg[r] = ok ? a.sx[i] : T(0);
        Cr[r] = act && rin[r] ? a.C[(int64_t)i * P + j] : T(0);
        lp |= rin[r] && jin && (nx[r] || ny);
    }
    lp = lp && act;
    int cidx[FRPT];
#pragma unroll
    for (int r = 0; r < FRPT; r++)
        cidx[r] = V == 1 && a.src.cell_idx && act && r < nr ? a.src.cell_idx[(int64_t)(ib + r) * P + j] : -1;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int s = g * a.Sg + wv; s < s1; s += NWAVES) {
        const T *__restrict__ w = a.w.at(s);
        const T *__restrict__ wo = a.wold.at(s);
        const T *__restrict__ phi = a.phi.at(s);
        const T *__restrict__ psi = a.psi.at(s);
        T W[FRPT + 2], WO[FRPT], PH[FRPT + 2], PS[FRPT], EL[FRPT], ER[FRPT], SL[FRPT], SR[FRPT];
#pragma unroll
        for (int q = 0; q < FRPT + 2; q++) {
            const int i = ib - 1 + q;
            const bool ok = act && q <= nr + 1 && i >= 0 && i <= NX + 1;
            W[q] = ok ? w[(int64_t)i * P + j] : T(0);
            PH[q] = ok && lp ? phi[(int64_t)i * P + j] : T(0);
        }
#pragma unroll
        for (int r = 0; r < FRPT; r++) {
            const int64_t row = (int64_t)(ib + r) * P;
            WO[r] = act && rin[r] ? wo[row + j] : T(0);
            PS[r] = lp && rin[r] ? psi[row + j] : T(0);
            EL[r] = le && rin[r] ? w[row + j - 1] : T(0);
            ER[r] = reok && rin[r] ? w[row + j + 1] : T(0);
            SL[r] = le && lp && rin[r] ? psi[row + j - 1] : T(0);
            SR[r] = reok && lp && rin[r] ? psi[row + j + 1] : T(0);
        }
        T *__restrict__ u = a.u.at(s);
#pragma unroll
        for (int r = 0; r < FRPT; r++) {
            if (r >= nr) break;  // uniform
            const int64_t row = (int64_t)(ib + r) * P;
            T wv1[1] = {W[r + 1]}, pv1[1] = {PS[r]};
            T ww, we, pw, pe;
            lr_of(wv1, EL[r], ER[r], lane, re, ww, we);
            lr_of(pv1, SL[r], SR[r], lane, re, pw, pe);
            const T wc = W[r + 1], wn = W[r], ws = W[r + 2];
            const T lap = k.cx * (ws + wn - T(2) * wc) + k.cy * (we + ww - T(2) * wc);
            const T s_ = sg[r], Cij = Cr[r];
            const bool in = rin[r] && jin;
            const bool near = nx[r] || ny;
            const T plain = T(2) * wc - WO[r] + Cij * lap;
            const T num = (T(2) - s_ * tu * k.dt2) * wc + Cij * lap + k.ex * (PH[r + 2] - PH[r]) +
                          k.ey * (pe - pw) - (T(1) - (s_ + tu) * k.hdt) * WO[r];
            T out = in ? (near ? num / (T(1) + (s_ + tu) * k.hdt) : plain) : T(0);
            if (V == 1 && cidx[r] >= 0) out = add_sources(a.src, cidx[r], s, out, k.dt2);
            if (act) u[row + j] = out;
            if (V == 1 && lp) {
                const bool pml = in && (s_ != T(0) || tu != T(0));
                a.phio.at(s)[row + j] = pml ? (T(1) - k.dt * s_) * PH[r + 1] + Cij * (tu - s_) * k.fx * (ws - wn) : T(0);
                a.psio.at(s)[row + j] = pml ? (T(1) - k.dt * tu) * PS[r] + Cij * (s_ - tu) * k.fy * (we - ww) : T(0);
            }
        }
    }
}

// Rows per item of the two-level frame work: 4, 8, .. 64 (FRAME2_H item lists).  Taller items
// read fewer halo rows, but each wavefront walks its item, h + 2 rows, once per shot of its
// quarter of the group, one shot after another.  With few frame wavefronts beside the
// rectangle's work that walk is the launch's critical path, about 2.3 us per row and shot for
// the two levels (32-row items at 1000^2 x 32 shots: 156 us per level against 98 us with
// 8-row items; at 2000^2 x 64 shots 64-row items are the fastest, 610 against 634 us).  The
// launch takes the tallest height not above the number of active shots whose walk fits in
// half the pair's streaming time at 5 TB/s.  per_wave: shots each wavefront walks;
// pair_bytes: algorithmic bytes of the two levels.
constexpr int FRAME2_H = 5;
inline int frame2_height(int k) { return 4 << k; }
inline int frame2_pick(int act, int per_wave, double pair_bytes)
{
    const double budget_us = 0.5 * pair_bytes / 5.0e6;
    int k = 0;
    while (k + 1 < FRAME2_H && frame2_height(k + 1) <= act &&
           2.3 * (frame2_height(k + 1) + 2) * per_wave <= budget_us)
        k++;
    return k;
}

// Two-level frame work (kernel 1, sources from the per-cell table), streaming form.  An item is
// nr rows of at most WAVE - 2 columns [m.z, m.w) of the complement of the rectangle shrunk by
// one cell; lane L holds column m.z - 1 + L, the first and last lanes supplying the halo
// columns.  For each shot the wavefront walks the level-n rows x = ib-1 .. ib+nr once, holding
// u^{n-1} and phi^{n-1} of rows x-1 .. x+1 and level n of rows x-2 .. x in registers, with the
// next row's inputs loaded one step ahead: every lane computes level n at its column (from
// levels n-1, n-2 and phi/psi^{n-1}); once level n of rows y-1 .. y+1 is known (y = x-1) it
// stores level n on row y (outside R: the rectangle pass stores it inside) and lanes 1 ..
// m.w - m.z compute level n+1 there.  Every value is fwd_frame_body's expression in the same
// order, so the results equal single-level stepping bit for bit.
template <typename T>
__device__ inline void fwd_frame2s_body(const FwdArgs<T> &a, const Lv<T> &u2, const Lv<T> &phi2,
                                        const Lv<T> &psi2, const Rect &R, const TileMap &map,
                                        const int4 *__restrict__ items, int nitems, int bid)
{
    int g, blk, tj;
    if (!decode_tile(map, g, blk, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int it = blk;
    if (it >= nitems) return;  // workgroup-uniform
    const int4 m = items[it];
    const int ib = m.x, nr = m.y, j = m.z - 1 + lane;
    const int NX = a.NX, NY = a.NY, P = a.P;
    const int jlast = min(m.w, NY + 1);  // the block's last column
    const bool act = j >= 0 && j <= jlast;
    const bool own = act && lane >= 1 && j < m.w;  // a column of the item
    const bool le = lane == 0 && act && j >= 1;
    const bool re = act && j == jlast;
    const bool reok = re && j + 1 <= NY + 1;
    const AcCoef<T> k = a.k;
    const bool jin = j >= 1 && j <= NY;
    const T tu = act ? a.ty[j] : T(0);
    const bool ny = act && a.neary[j] != 0;
    const bool jR = j >= R.jL && j < R.jR;
    // near cells on the block's rows ib-1 .. ib+nr (lp) and on the item's rows (lpi): some row
    // is interior (rb, ri) / interior and near in x (nb, ni)
    bool rb = false, nb = false, ri = false, ni = false;
    for (int x = max(ib - 1, 0); x <= min(ib + nr, NX + 1); x++) {  // uniform
        const bool rin = x >= 1 && x <= NX;
        const bool nxx = a.nearx[x] != 0;
        rb = rb || rin;
        nb = nb || (rin && nxx);
        if (x >= ib && x < ib + nr) {
            ri = ri || rin;
            ni = ni || (rin && nxx);
        }
    }
    const bool lp = act && jin && ((ny && rb) || nb);
    const bool lpi = act && jin && ((ny && ri) || ni);
    SrcIn<T> src2 = a.src;
    src2.n = a.src.n + 1;
    // inputs of level n on row x besides u^{n-1} and phi^{n-1}
    struct RowIn {
        T wo, ps, el, er, sl, sr, c, sg;
        int ci;
        bool lvl, rin, nx;
    };
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int s = g * a.Sg + wv; s < s1; s += NWAVES) {
        const T *__restrict__ w = a.w.at(s);
        const T *__restrict__ wo = a.wold.at(s);
        const T *__restrict__ phi = a.phi.at(s);
        const T *__restrict__ psi = a.psi.at(s);
        T *__restrict__ u = a.u.at(s);
        T *__restrict__ uo2 = u2.at(s);
        // u^{n-1} and phi^{n-1} of row x (rows ib-2 .. ib+nr+1 inside [0, NX+1])
        auto ldW = [&](int x, T &wx, T &px) {
            const bool ok = act && x >= 0 && x <= NX + 1 && x <= ib + nr + 1;
            wx = ok ? w[(int64_t)x * P + j] : T(0);
            px = ok && lp ? phi[(int64_t)x * P + j] : T(0);
        };
        auto ldR = [&](int x, RowIn &o) {
            o.lvl = x >= 0 && x <= NX + 1;
            o.rin = o.lvl && x >= 1 && x <= NX;
            o.nx = o.lvl && a.nearx[x] != 0;
            o.sg = o.lvl ? a.sx[x] : T(0);
            const int64_t row = (int64_t)x * P;
            o.c = act && o.rin ? a.C[row + j] : T(0);
            o.ci = act && o.lvl && a.src.cell_idx ? a.src.cell_idx[row + j] : -1;
            o.wo = act && o.rin ? wo[row + j] : T(0);
            o.ps = lp && o.rin ? psi[row + j] : T(0);
            o.el = le && o.rin ? w[row + j - 1] : T(0);
            o.er = reok && o.rin ? w[row + j + 1] : T(0);
            o.sl = le && lp && o.rin ? psi[row + j - 1] : T(0);
            o.sr = reok && lp && o.rin ? psi[row + j + 1] : T(0);
        };
        // window: u^{n-1} / phi^{n-1} of rows x-1 (w0, p0), x (w1, p1), x+1 (w2, p2); level n of
        // rows x-2 (um, hm) and x-1 (uc, hc, sc); rc = inputs of row x, ry = those of row x-1
        T w0, w1, w2, p0, p1, p2;
        ldW(ib - 2, w0, p0);
        ldW(ib - 1, w1, p1);
        ldW(ib, w2, p2);
        RowIn rc, ry;
        ldR(ib - 1, rc);
        ry = rc;
        T um = T(0), uc = T(0), hm = T(0), hc = T(0), sc = T(0);
        for (int x = ib - 1; x <= ib + nr; x++) {  // uniform
            // next row's inputs first, so their latency overlaps this step
            T w3 = T(0), p3 = T(0);
            RowIn rn = rc;
            if (x + 2 <= ib + nr + 1) ldW(x + 2, w3, p3);
            if (x + 1 <= ib + nr) ldR(x + 1, rn);
            // level n on row x
            T U, H, S;
            {
                T wv1[1] = {w1}, pv1[1] = {rc.ps};
                T ww, we, pw, pe;
                lr_of(wv1, rc.el, rc.er, lane, re, ww, we);
                lr_of(pv1, rc.sl, rc.sr, lane, re, pw, pe);
                const T wc = w1, wn = w0, ws = w2;
                const T lap = k.cx * (ws + wn - T(2) * wc) + k.cy * (we + ww - T(2) * wc);
                const T s_ = rc.sg, Cij = rc.c;
                const bool in = rc.rin && jin;
                const bool near = rc.nx || ny;
                const T plain = T(2) * wc - rc.wo + Cij * lap;
                const T num = (T(2) - s_ * tu * k.dt2) * wc + Cij * lap + k.ex * (p2 - p0) +
                              k.ey * (pe - pw) - (T(1) - (s_ + tu) * k.hdt) * rc.wo;
                T out = in ? (near ? num / (T(1) + (s_ + tu) * k.hdt) : plain) : T(0);
                if (rc.ci >= 0) out = add_sources(a.src, rc.ci, s, out, k.dt2);
                U = act && rc.lvl ? out : T(0);
                const bool pml = in && (s_ != T(0) || tu != T(0));
                H = lp && pml ? (T(1) - k.dt * s_) * p1 + Cij * (tu - s_) * k.fx * (ws - wn) : T(0);
                S = lp && pml ? (T(1) - k.dt * tu) * rc.ps + Cij * (s_ - tu) * k.fy * (we - ww) : T(0);
            }
            if (x >= ib + 1) {  // uniform: level n of rows y-1 .. y+1 is known
                const int y = x - 1;
                const int64_t row = (int64_t)y * P;
                if (own && !(jR && y >= R.iL && y < R.iR)) {
                    u[row + j] = uc;
                    if (lpi) {
                        a.phio.at(s)[row + j] = hc;
                        a.psio.at(s)[row + j] = sc;
                    }
                }
                T uv1[1] = {uc}, pv1[1] = {sc};
                T ww, we, pw, pe;
                lr_of(uv1, T(0), T(0), lane, false, ww, we);
                lr_of(pv1, T(0), T(0), lane, false, pw, pe);
                const T wc = uc, wn = um, ws = U;
                const T lap = k.cx * (ws + wn - T(2) * wc) + k.cy * (we + ww - T(2) * wc);
                const T s_ = ry.sg, Cij = ry.c;
                const bool in = ry.rin && jin;
                const bool near = ry.nx || ny;
                const T plain = T(2) * wc - w0 + Cij * lap;
                const T num = (T(2) - s_ * tu * k.dt2) * wc + Cij * lap + k.ex * (H - hm) +
                              k.ey * (pe - pw) - (T(1) - (s_ + tu) * k.hdt) * w0;
                T out = in ? (near ? num / (T(1) + (s_ + tu) * k.hdt) : plain) : T(0);
                if (ry.ci >= 0) out = add_sources(src2, ry.ci, s, out, k.dt2);
                if (own) uo2[row + j] = out;
                if (own && lpi) {
                    const bool pml = in && (s_ != T(0) || tu != T(0));
                    phi2.at(s)[row + j] = pml ? (T(1) - k.dt * s_) * hc + Cij * (tu - s_) * k.fx * (ws - wn) : T(0);
                    psi2.at(s)[row + j] = pml ? (T(1) - k.dt * tu) * sc + Cij * (s_ - tu) * k.fy * (we - ww) : T(0);
                }
            }
            // rotate
            um = uc;
            uc = U;
            hm = hc;
            hc = H;
            sc = S;
            w0 = w1;
            w1 = w2;
            w2 = w3;
            p0 = p1;
            p1 = p2;
            p2 = p3;
            ry = rc;
            rc = rn;
        }
    }
}

// Frame work of kernel 0 with its phi/psi update and its sources (one level; used when the
// sources come from the per-cell table).  Items of at most WAVE - 2 columns [m.z, m.w); lane L
// holds column m.z - 1 + L.  Every lane computes u^n before the sources at its column on the
// item's rows and one row above and below; the item's lanes then update phi/psi from those
// values (the expressions of ac_fwd_phipsi) and store u^n with the sources added (those of
// ac_src_rcv), so neither pass is launched.
template <typename T>
__device__ inline void fwd_frame0h_body(const FwdArgs<T> &a, const TileMap &map,
                                        const int4 *__restrict__ items, int nitems, int bid)
{
    constexpr int NB = FRPT + 2;
    int g, blk, tj;
    if (!decode_tile(map, g, blk, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int it = blk;
    if (it >= nitems) return;  // workgroup-uniform
    const int4 m = items[it];
    const int ib = m.x, nr = m.y, j = m.z - 1 + lane;
    const int NX = a.NX, NY = a.NY, P = a.P;
    const int jlast = min(m.w, NY + 1);
    const bool act = j >= 0 && j <= jlast;
    const bool own = act && lane >= 1 && j < m.w;
    const bool le = lane == 0 && act && j >= 1;
    const bool re = act && j == jlast;
    const bool reok = re && j + 1 <= NY + 1;
    const AcCoef<T> k = a.k;
    const bool jin = j >= 1 && j <= NY;
    const T tu = act ? a.ty[j] : T(0);
    const bool ny = act && a.neary[j] != 0;
    bool rin[NB], nx[NB], lvl[NB];
    T sg[NB], Cr[NB];
    int cidx[NB];
    bool lp = false, lpi = false;
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int i = ib - 1 + q;
        lvl[q] = q <= nr + 1 && i >= 0 && i <= NX + 1;
        rin[q] = lvl[q] && i >= 1 && i <= NX;
        nx[q] = lvl[q] && a.nearx[i] != 0;
        sg[q] = lvl[q] ? a.sx[i] : T(0);
        Cr[q] = act && rin[q] ? a.C[(int64_t)i * P + j] : T(0);
        cidx[q] = own && q >= 1 && q <= nr && a.src.cell_idx ? a.src.cell_idx[(int64_t)i * P + j] : -1;
        const bool nr_ = rin[q] && jin && (nx[q] || ny);
        lp |= nr_;
        if (q >= 1 && q <= nr) lpi |= nr_;
    }
    lp = lp && act;
    lpi = lpi && act;
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int s = g * a.Sg + wv; s < s1; s += NWAVES) {
        const T *__restrict__ w = a.w.at(s);
        const T *__restrict__ wo = a.wold.at(s);
        const T *__restrict__ phi = a.phi.at(s);
        const T *__restrict__ psi = a.psi.at(s);
        T W[NB + 2], PH[NB + 2], WO[NB], PS[NB], EL[NB], ER[NB], SL[NB], SR[NB];
#pragma unroll
        for (int q = 0; q < NB + 2; q++) {
            const int i = ib - 2 + q;
            const bool ok = act && q <= nr + 3 && i >= 0 && i <= NX + 1;
            W[q] = ok ? w[(int64_t)i * P + j] : T(0);
            PH[q] = ok && lp ? phi[(int64_t)i * P + j] : T(0);
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int64_t row = (int64_t)(ib - 1 + q) * P;
            WO[q] = act && rin[q] ? wo[row + j] : T(0);
            PS[q] = lp && rin[q] ? psi[row + j] : T(0);
            EL[q] = le && rin[q] ? w[row + j - 1] : T(0);
            ER[q] = reok && rin[q] ? w[row + j + 1] : T(0);
            SL[q] = le && lp && rin[q] ? psi[row + j - 1] : T(0);
            SR[q] = reok && lp && rin[q] ? psi[row + j + 1] : T(0);
        }
        T Un[NB];  // u^n before the sources
#pragma unroll
        for (int q = 0; q < NB; q++) {
            T wv1[1] = {W[q + 1]}, pv1[1] = {PS[q]};
            T ww, we, pw, pe;
            lr_of(wv1, EL[q], ER[q], lane, re, ww, we);
            lr_of(pv1, SL[q], SR[q], lane, re, pw, pe);
            const T wc = W[q + 1], wn = W[q], ws = W[q + 2];
            const T lap = k.cx * (ws + wn - T(2) * wc) + k.cy * (we + ww - T(2) * wc);
            const T s_ = sg[q], Cij = Cr[q];
            const bool in = rin[q] && jin;
            const bool near = nx[q] || ny;
            const T plain = T(2) * wc - WO[q] + Cij * lap;
            const T num = (T(2) - s_ * tu * k.dt2) * wc + Cij * lap + k.ex * (PH[q + 2] - PH[q]) +
                          k.ey * (pe - pw) - (T(1) - (s_ + tu) * k.hdt) * WO[q];
            const T out = in ? (near ? num / (T(1) + (s_ + tu) * k.hdt) : plain) : T(0);
            Un[q] = act && lvl[q] ? out : T(0);
        }
        T *__restrict__ u = a.u.at(s);
#pragma unroll
        for (int r = 0; r < FRPT; r++) {
            if (r >= nr) break;  // uniform
            const int q = r + 1, i = ib + r;
            const int64_t ij = (int64_t)i * P + j;
            T uv1[1] = {Un[q]};
            T uw, ue;
            lr_of(uv1, T(0), T(0), lane, false, uw, ue);
            if (own && lpi) {
                const T s_ = sg[q], Cij = Cr[q];
                const bool pml = rin[q] && jin && (s_ != T(0) || tu != T(0));
                a.phio.at(s)[ij] = pml ? (T(1) - k.dt * s_) * PH[q + 1] + Cij * (tu - s_) * k.fx * (Un[q + 1] - Un[q - 1]) : T(0);
                a.psio.at(s)[ij] = pml ? (T(1) - k.dt * tu) * PS[q] + Cij * (s_ - tu) * k.fy * (ue - uw) : T(0);
            }
            T out = Un[q];
            if (cidx[q] >= 0) out = add_sources(a.src, cidx[q], s, out, k.dt2);
            if (own) u[ij] = out;
        }
    }
}

// The frame work of a level: V = 0, 1 the PropagatorKernel; V = 2 kernel 0 with its phi/psi
// update and sources in the frame work (fwd_frame0h_body, items of WAVE - 2 columns).
template <typename T, int V>
__device__ inline void frame_work(const FwdArgs<T> &a, const TileMap &map, const int4 *__restrict__ items,
                                  int nitems, int bid)
{
    if constexpr (V == 2) fwd_frame0h_body<T>(a, map, items, nitems, bid);
    else fwd_frame_body<T, V>(a, map, items, nitems, bid);
}

// kernel 0: phi^n/psi^n from the new u (before sources), PML cells only.  The PML lies in the
// frame: one workgroup per (frame item, 4 shots), a wavefront per shot, a lane per column.
template <typename T>
__global__ __launch_bounds__(256) void ac_fwd_phipsi(
    Lv<const T> u_, Lv<const T> phi_, Lv<const T> psi_, Lv<T> phio, Lv<T> psio,
    const T *__restrict__ C, const T *__restrict__ sx, const T *__restrict__ ty,
    const uint8_t *__restrict__ pmlx, const uint8_t *__restrict__ pmly, int NX, int NY, int P,
    AcCoef<T> k, const int4 *__restrict__ items, int act)
{
    const int4 m = items[blockIdx.x];
    const int j = m.z + (threadIdx.x & 63);
    const int s = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (s >= act || j >= m.w || j < 1 || j > NY) return;
    const T *__restrict__ u = u_.at(s);
    const T tu = ty[j];
    for (int r = 0; r < m.y; r++) {
        const int i = m.x + r;
        if (i < 1 || i > NX || !(pmlx[i] | pmly[j])) continue;
        const int64_t ij = (int64_t)i * P + j;
        const T sg = sx[i], Cij = C[ij];
        phio.at(s)[ij] = (T(1) - k.dt * sg) * phi_.at(s)[ij] + Cij * (tu - sg) * k.fx * (u[ij + P] - u[ij - P]);
        psio.at(s)[ij] = (T(1) - k.dt * tu) * psi_.at(s)[ij] + Cij * (sg - tu) * k.fy * (u[ij + 1] - u[ij - 1]);
    }
}

// Per-shot source injection at level n (u[cell] += dt^2 * srcv[n-1, q] in source order,
// Core.jl:593-601), then receiver sampling trace[n, r] = u[rcv_r] (Core.jl:726-730).
// One workgroup per shot; shots are in plan order.
template <typename T>
__global__ __launch_bounds__(256) void ac_src_rcv(
    Lv<T> u_, int n, const int64_t *__restrict__ us_ptr, const int64_t *__restrict__ us_cell,
    const int64_t *__restrict__ us_eptr, const int64_t *__restrict__ us_ent,
    const T *__restrict__ srcv, const int64_t *__restrict__ srcv_off,
    const int64_t *__restrict__ nsrc, T dt2, const int64_t *__restrict__ rcv_ptr,
    const int64_t *__restrict__ rcv_cell, T *__restrict__ traces,
    const int64_t *__restrict__ trace_off, int record)
{
    const int s = blockIdx.x;
    T *__restrict__ u = u_.at(s);
    const int64_t ns = nsrc[s];
    const T *__restrict__ row = srcv + srcv_off[s] + (int64_t)(n - 1) * ns;
    for (int64_t c = us_ptr[s] + threadIdx.x; c < us_ptr[s + 1]; c += blockDim.x) {
        const int64_t cell = us_cell[c];
        T acc = u[cell];
        for (int64_t e = us_eptr[c]; e < us_eptr[c + 1]; e++) acc += row[us_ent[e]] * dt2;
        u[cell] = acc;
    }
    if (!record) return;
    __syncthreads();
    const int64_t r0 = rcv_ptr[s], nr = rcv_ptr[s + 1] - r0;
    T *__restrict__ tr = traces + trace_off[s] + (int64_t)n * nr;
    for (int64_t r = threadIdx.x; r < nr; r += blockDim.x) tr[r] = u[rcv_cell[r0 + r]];
}

// Receiver sampling of a whole forward segment at once, trace[n, r] = u^n[rcv_r] for levels
// lo < n <= hi, from the checkpoint slot ring (level n in slot n mod NS; sources already added).
// Used when the frame work adds the sources.  Grid: (levels, shots).
template <typename T>
__global__ __launch_bounds__(256) void ac_record(const T *__restrict__ seg, int64_t NS, int64_t slot_stride,
                                                 int64_t npad, int64_t lo, const int64_t *__restrict__ nstep,
                                                 const int64_t *__restrict__ rcv_ptr,
                                                 const int64_t *__restrict__ rcv_cell, T *__restrict__ traces,
                                                 const int64_t *__restrict__ trace_off)
{
    const int64_t n = lo + 1 + blockIdx.x;
    const int s = blockIdx.y;
    if (nstep[s] < n) return;
    const T *__restrict__ u = seg + (n % NS) * slot_stride + (int64_t)s * npad;
    const int64_t r0 = rcv_ptr[s], nr = rcv_ptr[s + 1] - r0;
    T *__restrict__ tr = traces + trace_off[s] + n * nr;
    for (int64_t r = threadIdx.x; r < nr; r += blockDim.x) tr[r] = u[rcv_cell[r0 + r]];
}

// ---------------------------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------------------------
// kernel 0 only: F^n, G^n on PML cells and their dL/dC term with v = u^n before sources.  The
// PML lies in the frame: one workgroup per (frame item, shot group); its 4 wavefronts take every
// 4th shot of the group, a lane per column, and their dL/dC partial sums are added in wavefront
// order through LDS into one write per cell.
template <typename T>
__global__ __launch_bounds__(256) void ac_adj_phipsi(
    Lv<const T> P1_, Lv<const T> F1_, Lv<const T> G1_, Lv<T> F0_, Lv<T> G0_, Lv<const T> un_,
    Lv<T> gC_, int Sg, int act, int n, const T *__restrict__ sx, const T *__restrict__ ty,
    const uint8_t *__restrict__ pmlx, const uint8_t *__restrict__ pmly,
    const int64_t *__restrict__ us_ptr, const int64_t *__restrict__ us_cell,
    const int64_t *__restrict__ us_eptr, const int64_t *__restrict__ us_ent,
    const T *__restrict__ srcv, const int64_t *__restrict__ srcv_off,
    const int64_t *__restrict__ nsrc, int NX, int NY, int P, AcCoef<T> k,
    const int4 *__restrict__ items)
{
    const int4 m = items[blockIdx.x];
    const int g = blockIdx.y;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = m.z + lane;
    const bool jok = j < m.w && j >= 1 && j <= NY;
    const T tu = jok ? ty[j] : T(0);
    T acc[FRPT] = {};
    bool cell[FRPT];
#pragma unroll
    for (int r = 0; r < FRPT; r++) {
        const int i = m.x + r;
        cell[r] = jok && r < m.y && i >= 1 && i <= NX && (pmlx[i] | pmly[j]);
    }
    const int s1 = min(act, (g + 1) * Sg);
    for (int s = g * Sg + wv; s < s1; s += NWAVES) {
        const T *__restrict__ P1 = P1_.at(s);
        const T *__restrict__ u = un_.at(s);
        const int64_t ns = nsrc[s];
        const T *__restrict__ row = srcv + srcv_off[s] + (int64_t)(n - 1) * ns;
#pragma unroll
        for (int r = 0; r < FRPT; r++) {
            if (!cell[r]) continue;
            const int i = m.x + r;
            const int64_t ij = (int64_t)i * P + j;
            const T sg = sx[i];
            const T F0 = (T(1) - k.dt * sg) * F1_.at(s)[ij] + k.ex * (P1[ij - P] - P1[ij + P]);
            const T G0 = (T(1) - k.dt * tu) * G1_.at(s)[ij] + k.ey * (P1[ij - 1] - P1[ij + 1]);
            F0_.at(s)[ij] = F0;
            G0_.at(s)[ij] = G0;
            T vn = u[ij - P], vs = u[ij + P], vw = u[ij - 1], ve = u[ij + 1];
            // remove this level's sources from the neighbours (v is u^n before injection)
            for (int64_t c = us_ptr[s]; c < us_ptr[s + 1]; c++) {
                const int64_t cl = us_cell[c];
                if (cl != ij - P && cl != ij + P && cl != ij - 1 && cl != ij + 1) continue;
                T add = T(0);
                for (int64_t e = us_eptr[c]; e < us_eptr[c + 1]; e++) add += row[us_ent[e]] * k.dt2;
                if (cl == ij - P) vn -= add;
                if (cl == ij + P) vs -= add;
                if (cl == ij - 1) vw -= add;
                if (cl == ij + 1) ve -= add;
            }
            acc[r] += F0 * (tu - sg) * k.fx * (vs - vn) + G0 * (sg - tu) * k.fy * (ve - vw);
        }
    }
    __shared__ T red[NWAVES][FRPT][WAVE];
#pragma unroll
    for (int r = 0; r < FRPT; r++) red[wv][r][lane] = acc[r];
    __syncthreads();
    if (wv != 0) return;
#pragma unroll
    for (int r = 0; r < FRPT; r++) {
        if (!cell[r]) continue;
        T sum = red[0][r][lane];
#pragma unroll
        for (int q = 1; q < NWAVES; q++) sum += red[q][r][lane];
        gC_.at(g)[(int64_t)(m.x + r) * P + j] += sum;
    }
}

template <typename T> struct AdjArgs {
    Lv<T> Pn;                 // in: P^{n+2}, out: P^n
    Lv<const T> P1, Fn, Gn;   // P^{n+1}; F/G read at neighbours (F^{n+1} kernel 1, F^n kernel 0)
    Lv<T> F0, G0;             // kernel 1: F^n, G^n
    Lv<const T> um1, dense;   // u^{n-1}; optional dense cotangent of u^n
    Lv<T> gC;                 // per shot group
    const T *C, *sx, *ty;
    const uint8_t *nearx, *neary;
    int NX, NY, P, act, Sg;
    TileMap map;
    AcCoef<T> k;
};

// Rectangle adjoint: P^n = 2 P^{n+1} + lap(C P^{n+1}) - P^{n+2} (+ the dense cotangent of u^n
// in the full-history mode), dL/dC += P^n lap(u^{n-1}); no PML term reaches these cells.
template <typename T, int RPT, int VEC, int SB>
__device__ inline void adj_rect_body(const AdjArgs<T> &a, const Rect &R, int bid)
{
    constexpr bool DENSE = VEC == 1;  // the dense cotangent only exists in the history layout
    int g, strip, tj;
    if (!decode_tile(a.map, g, strip, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int j0 = R.jL + (tj * WAVE + lane) * VEC;
    const int ib = R.iL + (strip * NWAVES + wv) * RPT;
    if (ib >= R.iR) return;  // wave-uniform
    const int P = a.P, nr = min(RPT, R.iR - ib);
    const bool act = j0 < R.jR;
    const bool le = lane == 0 && act;
    const bool re = act && (lane == WAVE - 1 || j0 + VEC == R.jR);
    const bool dense = DENSE && a.dense.p != nullptr;
    const AcCoef<T> k = a.k;
    T Cr[RPT + 2][VEC], cl[RPT], cr[RPT];
    // byte offsets: row x starts at x * pb, this lane's vector at jo
    const uint32_t pb = (uint32_t)P * (uint32_t)sizeof(T), jo = (uint32_t)j0 * (uint32_t)sizeof(T);
#pragma unroll
    for (int q = 0; q < RPT + 2; q++) {
        if (act && q <= nr + 1) vld(Cr[q], at_off(a.C, (uint32_t)(ib - 1 + q) * pb + jo));
        else vzero(Cr[q]);
    }
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int64_t row = (int64_t)(ib + r) * P;
        const T el = le && r < nr ? a.C[row + j0 - 1] : T(0);
        const T er = re && r < nr ? a.C[row + j0 + VEC] : T(0);
        lr_of(Cr[r + 1], el, er, lane, re, cl[r], cr[r]);
    }
    T gacc[RPT][VEC];
#pragma unroll
    for (int r = 0; r < RPT; r++) vzero(gacc[r]);
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int s = g * a.Sg; s < s1; s += SB) {
        T P1[SB][RPT + 2][VEC], U[SB][RPT + 2][VEC], P2[SB][RPT][VEC], DN[SB][RPT][VEC];
        T PL[SB][RPT], PR[SB][RPT], UL[SB][RPT], UR[SB][RPT];
#pragma unroll
        for (int b = 0; b < SB; b++) {
            if (s + b >= s1) break;  // uniform
            const T *__restrict__ p1 = a.P1.at(s + b);
            const T *__restrict__ um = a.um1.at(s + b);
            const T *__restrict__ p2 = a.Pn.at(s + b);
#pragma unroll
            for (int q = 0; q < RPT + 2; q++) {
                const uint32_t row = (uint32_t)(ib - 1 + q) * pb + jo;
                if (act && q <= nr + 1) {
                    vld(P1[b][q], at_off(p1, row));
                    vld(U[b][q], at_off(um, row));
                } else {
                    vzero(P1[b][q]);
                    vzero(U[b][q]);
                }
            }
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                const uint32_t ro = (uint32_t)(ib + r) * pb + jo;
                const bool ok = act && r < nr;
                if (ok) vld(P2[b][r], at_off(p2, ro));
                else vzero(P2[b][r]);
                if (dense) {
                    if (ok) vld(DN[b][r], at_off(a.dense.at(s + b), ro));
                    else vzero(DN[b][r]);
                }
                PL[b][r] = le && r < nr ? at_off(p1, ro)[-1] : T(0);
                PR[b][r] = re && r < nr ? at_off(p1, ro)[VEC] : T(0);
                UL[b][r] = le && r < nr ? at_off(um, ro)[-1] : T(0);
                UR[b][r] = re && r < nr ? at_off(um, ro)[VEC] : T(0);
            }
        }
#pragma unroll
        for (int b = 0; b < SB; b++) {
            if (s + b >= s1) break;
            T *__restrict__ Pout = a.Pn.at(s + b);
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                if (r >= nr) break;  // uniform
                T pl, pr, ul, ur;
                lr_of(P1[b][r + 1], PL[b][r], PR[b][r], lane, re, pl, pr);
                lr_of(U[b][r + 1], UL[b][r], UR[b][r], lane, re, ul, ur);
                T out[VEC];
#pragma unroll
                for (int e = 0; e < VEC; e++) {
                    const T p1c = P1[b][r + 1][e];
                    const T p1w = e == 0 ? pl : P1[b][r + 1][e - 1];
                    const T p1e = e == VEC - 1 ? pr : P1[b][r + 1][e + 1];
                    const T Cw = e == 0 ? cl[r] : Cr[r + 1][e - 1];
                    const T Ce = e == VEC - 1 ? cr[r] : Cr[r + 1][e + 1];
                    const T qc = Cr[r + 1][e] * p1c;
                    const T lapq = k.cx * (Cr[r + 2][e] * P1[b][r + 2][e] + Cr[r][e] * P1[b][r][e] - T(2) * qc) +
                                   k.cy * (Ce * p1e + Cw * p1w - T(2) * qc);
                    T Pv = T(2) * p1c + lapq - P2[b][r][e];
                    if (dense) Pv += DN[b][r][e];
                    out[e] = Pv;
                    const T uc = U[b][r + 1][e];
                    const T uw = e == 0 ? ul : U[b][r + 1][e - 1];
                    const T ue = e == VEC - 1 ? ur : U[b][r + 1][e + 1];
                    gacc[r][e] += Pv * (k.cx * (U[b][r + 2][e] + U[b][r][e] - T(2) * uc) + k.cy * (ue + uw - T(2) * uc));
                }
                if (act) vst(at_off(Pout, (uint32_t)(ib + r) * pb + jo), out);
            }
        }
    }
    if (!act) return;
    T *__restrict__ gC = a.gC.at(g);
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        if (r >= nr) break;
        T *p = at_off(gC, (uint32_t)(ib + r) * pb + jo);
        T cur[VEC];
        vld(cur, p);
#pragma unroll
        for (int e = 0; e < VEC; e++) cur[e] += gacc[r][e];
        vst(p, cur);
    }
}

// Frame adjoint, one depth cell per lane: the full PML transpose (F/G read at neighbours,
// F^n/G^n written for kernel 1 on PML lanes), zero on the padding ring.  Work split as in the
// frame forward; the wavefronts' dL/dC partial sums are combined through LDS.
template <typename T, int V>
__device__ inline void adj_frame_body(const AdjArgs<T> &a, const TileMap &map,
                                      const int4 *__restrict__ items, int nitems, int bid)
{
    int g, blk, tj;
    if (!decode_tile(map, g, blk, tj, bid)) return;
    const int lane = threadIdx.x & (WAVE - 1), wv = wave_id();
    const int it = blk;  // one item per workgroup; its 4 wavefronts take every 4th shot
    if (it >= nitems) return;  // workgroup-uniform
    const int4 m = items[it];
    const int ib = m.x, nr = m.y, j = m.z + lane;
    const int NX = a.NX, NY = a.NY, P = a.P;
    const bool act = j < m.w;
    const bool le = lane == 0 && act && j >= 1;
    const bool re = act && (lane == WAVE - 1 || j + 1 == m.w);
    const bool reok = re && j + 1 <= NY + 1;
    const AcCoef<T> k = a.k;
    const bool jin = j >= 1 && j <= NY;
    const T tu = act ? a.ty[j] : T(0);
    const T tw = act && j >= 1 ? a.ty[j - 1] : T(0);
    const T te = act && j + 1 <= NY + 1 ? a.ty[j + 1] : T(0);
    const bool ny = act && a.neary[j] != 0;
    bool rin[FRPT], nx[FRPT];
    T sg[FRPT], sn[FRPT], ss[FRPT], Cr[FRPT + 2], cl[FRPT], cr[FRPT];
    bool lp = false;
#pragma unroll
    for (int q = 0; q < FRPT + 2; q++) {
        const int i = ib - 1 + q;
        Cr[q] = act && q <= nr + 1 && i >= 0 && i <= NX + 1 ? a.C[(int64_t)i * P + j] : T(0);
    }
#pragma unroll
    for (int r = 0; r < FRPT; r++) {
        const int i = ib + r;
        const bool ok = r < nr;
        rin[r] = ok && i >= 1 && i <= NX;
        nx[r] = ok && a.nearx[i] != 0;
        sg[r] = ok ? a.sx[i] : T(0);
        sn[r] = rin[r] ? a.sx[i - 1] : T(0);
        ss[r] = rin[r] ? a.sx[i + 1] : T(0);
        lp |= rin[r] && jin && (nx[r] || ny);
        const T el = le && rin[r] ? a.C[(int64_t)i * P + j - 1] : T(0);
        const T er = reok && rin[r] ? a.C[(int64_t)i * P + j + 1] : T(0);
        T c1[1] = {Cr[r + 1]};
        lr_of(c1, el, er, lane, re, cl[r], cr[r]);
    }
    lp = lp && act;
    const bool dense = a.dense.p != nullptr;
    T gacc[FRPT] = {};
    const int s1 = min(a.act, (g + 1) * a.Sg);
    for (int s = g * a.Sg + wv; s < s1; s += NWAVES) {
        const T *__restrict__ p1 = a.P1.at(s);
        const T *__restrict__ um = a.um1.at(s);
        const T *__restrict__ Fn = a.Fn.at(s);
        const T *__restrict__ Gn = a.Gn.at(s);
        T *__restrict__ Pout = a.Pn.at(s);
        T P1[FRPT + 2], U[FRPT + 2], F[FRPT + 2], P2[FRPT], DN[FRPT], G[FRPT];
        T PL[FRPT], PR[FRPT], UL[FRPT], UR[FRPT], GL[FRPT], GR[FRPT];
#pragma unroll
        for (int q = 0; q < FRPT + 2; q++) {
            const int i = ib - 1 + q;
            const bool ok = act && q <= nr + 1 && i >= 0 && i <= NX + 1;
            const int64_t ij = (int64_t)i * P + j;
            P1[q] = ok ? p1[ij] : T(0);
            U[q] = ok ? um[ij] : T(0);
            F[q] = ok && lp ? Fn[ij] : T(0);
        }
#pragma unroll
        for (int r = 0; r < FRPT; r++) {
            const int64_t ij = (int64_t)(ib + r) * P + j;
            const bool ok = act && r < nr;
            P2[r] = ok ? Pout[ij] : T(0);
            DN[r] = ok && dense ? a.dense.at(s)[ij] : T(0);
            G[r] = lp && rin[r] ? Gn[ij] : T(0);
            PL[r] = le && rin[r] ? p1[ij - 1] : T(0);
            PR[r] = reok && rin[r] ? p1[ij + 1] : T(0);
            UL[r] = le && rin[r] ? um[ij - 1] : T(0);
            UR[r] = reok && rin[r] ? um[ij + 1] : T(0);
            GL[r] = le && lp && rin[r] ? Gn[ij - 1] : T(0);
            GR[r] = reok && lp && rin[r] ? Gn[ij + 1] : T(0);
        }
#pragma unroll
        for (int r = 0; r < FRPT; r++) {
            if (r >= nr) break;  // uniform
            const int64_t ij = (int64_t)(ib + r) * P + j;
            T v1[1] = {P1[r + 1]}, v2[1] = {U[r + 1]}, v3[1] = {G[r]};
            T p1w, p1e, uw, ue, gw, ge;
            lr_of(v1, PL[r], PR[r], lane, re, p1w, p1e);
            lr_of(v2, UL[r], UR[r], lane, re, uw, ue);
            lr_of(v3, GL[r], GR[r], lane, re, gw, ge);
            const T p1c = P1[r + 1], p1n = P1[r], p1s = P1[r + 2];
            const T Cc = Cr[r + 1], Cn = Cr[r], Cs = Cr[r + 2], Cw = cl[r], Ce = cr[r];
            const T uc = U[r + 1], un = U[r], us = U[r + 2];
            const T qc = Cc * p1c;
            const T lapq = k.cx * (Cs * p1s + Cn * p1n - T(2) * qc) + k.cy * (Ce * p1e + Cw * p1w - T(2) * qc);
            const T s_ = sg[r], p2 = P2[r];
            const bool in = rin[r] && jin;
            const bool near = nx[r] || ny;
            T Pv;
            if (near) {
                T U_ = (T(2) - s_ * tu * k.dt2) * p1c + lapq - (T(1) - (s_ + tu) * k.hdt) * p2 +
                       k.fx * (Cn * (tu - sn[r]) * F[r] - Cs * (tu - ss[r]) * F[r + 2]) +
                       k.fy * (Cw * (s_ - tw) * gw - Ce * (s_ - te) * ge);
                if (dense) U_ += DN[r];
                Pv = U_ / (T(1) + (s_ + tu) * k.hdt);
            } else {
                Pv = T(2) * p1c + lapq - p2;
                if (dense) Pv += DN[r];
            }
            Pv = in ? Pv : T(0);
            T gg = Pv * (k.cx * (us + un - T(2) * uc) + k.cy * (ue + uw - T(2) * uc));
            const bool pml = in && (s_ != T(0) || tu != T(0));
            T fo = T(0), go = T(0);
            if (V == 1 && pml) {
                fo = (T(1) - k.dt * s_) * F[r + 1] + k.ex * (p1n - p1s);
                go = (T(1) - k.dt * tu) * G[r] + k.ey * (p1w - p1e);
                gg += fo * (tu - s_) * k.fx * (us - un) + go * (s_ - tu) * k.fy * (ue - uw);
            }
            if (act) Pout[ij] = Pv;
            if (V == 1 && lp) {
                a.F0.at(s)[ij] = fo;
                a.G0.at(s)[ij] = go;
            }
            gacc[r] += in ? gg : T(0);
        }
    }
    // the group's dL/dC: the four wavefronts' partial sums, added in wavefront order
    __shared__ T red[NWAVES][FRPT][WAVE];
#pragma unroll
    for (int r = 0; r < FRPT; r++) red[wv][r][lane] = gacc[r];
    __syncthreads();
    if (wv != 0 || !act) return;
    T *__restrict__ gC = a.gC.at(g);
#pragma unroll
    for (int r = 0; r < FRPT; r++) {
        if (r >= nr) break;
        if (!rin[r] || !jin) continue;
        T sum = red[0][r][lane];
#pragma unroll
        for (int q = 1; q < NWAVES; q++) sum += red[q][r][lane];
        gC[(int64_t)(ib + r) * P + j] += sum;
    }
}

// Kernels.  ac_*_rect and ac_*_frame run one part of the grid each (the dense history layout
// uses them).  ac_*_level runs both in one launch: workgroups [0, nfb) take the frame items
// (tile map fmap), the rest the rectangle tiles (a.map), so the latency-bound frame work
// overlaps the bandwidth-bound rectangle work instead of following it.
template <typename T, int RPT, int VEC, int SB>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_fwd_rect(const FwdArgs<T> a, const Rect R)
{
    fwd_rect_body<T, RPT, VEC, SB>(a, R, (int)blockIdx.x);
}
template <typename T, int V>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_fwd_frame(const FwdArgs<T> a,
                                                              const int4 *__restrict__ items,
                                                              int nitems)
{
    frame_work<T, V>(a, a.map, items, nitems, (int)blockIdx.x);
}
template <typename T, int RPT, int VEC, int SB, int V>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_fwd_level(const FwdArgs<T> a, const Rect R,
                                                              const TileMap fmap,
                                                              const int4 *__restrict__ items,
                                                              int nitems, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) frame_work<T, V>(a, fmap, items, nitems, b);
    else fwd_rect_body<T, RPT, VEC, SB>(a, R, b - nfb);
}
// Levels n and n+1 of the rectangle (two-level body) with level n of the frame, in one launch.
template <typename T, int RPT, int VEC, int V>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_fwd_level2(const FwdArgs<T> a, const Lv<T> u2,
                                                               const Rect R, const TileMap fmap,
                                                               const int4 *__restrict__ items,
                                                               int nitems, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) frame_work<T, V>(a, fmap, items, nitems, b);
    else fwd_rect2_body<T, RPT, VEC>(a, u2, R, b - nfb);
}
// Levels n and n+1 of the whole grid in one launch: the two-level rectangle body and the
// two-level frame body over the complement of the shrunk rectangle (kernel 1).
template <typename T, int RPT, int VEC>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_fwd_pair(const FwdArgs<T> a, const Lv<T> u2,
                                                             const Lv<T> phi2, const Lv<T> psi2,
                                                             const Rect R, const TileMap fmap,
                                                             const int4 *__restrict__ items,
                                                             int nitems, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) fwd_frame2s_body<T>(a, u2, phi2, psi2, R, fmap, items, nitems, b);
    else fwd_rect2_body<T, RPT, VEC>(a, u2, R, b - nfb);
}
// Levels n and n+1 of the rectangle (streaming body) with level n of the frame (kernel 0).
template <typename T, int VEC, int ROWS, int V>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_fwd_level2_s(const FwdArgs<T> a, const Lv<T> u2,
                                                                 const Rect R, const TileMap fmap,
                                                                 const int4 *__restrict__ items,
                                                                 int nitems, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) frame_work<T, V>(a, fmap, items, nitems, b);
    else fwd_rect2s_body<T, VEC, ROWS>(a, u2, R, b - nfb);
}
// The same with the streaming rectangle body (ROWS rows per wavefront, one shot each).
template <typename T, int VEC, int ROWS>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_fwd_pair_s(const FwdArgs<T> a, const Lv<T> u2,
                                                               const Lv<T> phi2, const Lv<T> psi2,
                                                               const Rect R, const TileMap fmap,
                                                               const int4 *__restrict__ items,
                                                               int nitems, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) fwd_frame2s_body<T>(a, u2, phi2, psi2, R, fmap, items, nitems, b);
    else fwd_rect2s_body<T, VEC, ROWS>(a, u2, R, b - nfb);
}
template <typename T, int RPT, int VEC, int SB>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_adj_rect(const AdjArgs<T> a, const Rect R)
{
    adj_rect_body<T, RPT, VEC, SB>(a, R, (int)blockIdx.x);
}
template <typename T, int V>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_adj_frame(const AdjArgs<T> a,
                                                              const int4 *__restrict__ items,
                                                              int nitems)
{
    adj_frame_body<T, V>(a, a.map, items, nitems, (int)blockIdx.x);
}
template <typename T, int RPT, int VEC, int SB, int V>
__global__ __launch_bounds__(WAVE *NWAVES) void ac_adj_level(const AdjArgs<T> a, const Rect R,
                                                              const TileMap fmap,
                                                              const int4 *__restrict__ items,
                                                              int nitems, int nfb)
{
    const int b = (int)blockIdx.x;
    if (b < nfb) adj_frame_body<T, V>(a, fmap, items, nitems, b);
    else adj_rect_body<T, RPT, VEC, SB>(a, R, b - nfb);
}

// Receiver residual injection at level n after the stencil pass.  One wavefront owns each
// distinct interior receiver cell of a shot group; its lanes take the group's shots with
// receivers there (one segment of entries per shot): R = sum of that shot's residuals at the
// cell (in receiver order), P_s += R/d, and the group's dL/dC gets the sum over the shots of
// (R/d) lap(u_s^{n-1}), reduced across the lanes in a fixed order (no atomics).
template <typename T>
__global__ __launch_bounds__(256) void ac_adj_inject(
    Lv<T> P_, Lv<const T> um1_, Lv<T> gC_, int act, int n, const int64_t *__restrict__ gi_ptr,
    const int64_t *__restrict__ gi_cell, const int64_t *__restrict__ gi_sptr,
    const int32_t *__restrict__ gi_sshot, const int64_t *__restrict__ gi_seg,
    const int64_t *__restrict__ gi_rcv, const int64_t *__restrict__ rcv_ptr,
    const T *__restrict__ resid, const int64_t *__restrict__ trace_off,
    const T *__restrict__ sx, const T *__restrict__ ty, int NX, int NY, int P, AcCoef<T> k)
{
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t c = gi_ptr[g] + (int64_t)blockIdx.x * NWAVES + wave_id();
    if (c >= gi_ptr[g + 1]) return;  // wave-uniform
    const int64_t cell = gi_cell[c];
    const int i = (int)(cell / P), j = (int)(cell % P);
    if (i < 1 || i > NX || j < 1 || j > NY) return;  // ring cells carry no gradient
    const T den = T(1) + (sx[i] + ty[j]) * k.hdt;
    T acc = T(0);
    for (int64_t q = gi_sptr[c] + lane; q < gi_sptr[c + 1]; q += WAVE) {
        const int s = gi_sshot[q];
        if (s >= act) continue;
        const int64_t nr = rcv_ptr[s + 1] - rcv_ptr[s];
        const T *__restrict__ row = resid + trace_off[s] + (int64_t)n * nr;
        T R = T(0);
        for (int64_t e = gi_seg[q]; e < gi_seg[q + 1]; e++) R += row[gi_rcv[e]];
        const T add = R / den;
        P_.at(s)[cell] += add;
        const T *__restrict__ um = um1_.at(s);
        const T uc = um[cell];
        acc += add * (k.cx * (um[cell + P] + um[cell - P] - T(2) * uc) +
                      k.cy * (um[cell + 1] + um[cell - 1] - T(2) * uc));
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) gC_.at(g)[cell] += acc;
}

// model (caller layout, (NX+2) x (NY+2)) -> C (pitched), squared if asked
template <typename T>
__global__ void k_model_in(const T *__restrict__ m, T *__restrict__ C, int NX, int NY, int P, int sq)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= NY + 2) return;
    const T v = m[(int64_t)i * (NY + 2) + j];
    C[(int64_t)i * P + j] = sq ? v * v : v;
}

// sum of the shot-group accumulators (pitched) -> gradient in caller layout (x 2m if squared)
template <typename T>
__global__ void k_grad_out(const T *__restrict__ gC, int64_t stride, int ngroups,
                           const T *__restrict__ m, T *__restrict__ g, int NX, int NY, int P, int sq)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= NY + 2) return;
    const int64_t q = (int64_t)i * P + j, o = (int64_t)i * (NY + 2) + j;
    T acc = T(0);
    for (int s = 0; s < ngroups; s++) acc += gC[(int64_t)s * stride + q];
    g[o] = sq ? T(2) * m[o] * acc : acc;
}


// ---------------------------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------------------------
struct Shape {
    int rpt = 2, sg = 8, sb = 1;  // rows per wavefront, shots per workgroup, shots per batch
};

struct AcPlanBase {
    ads_acoustic_params p{};
    int dtype = ADS_F32;
    int variant = 0;  // 0 or 1
    int64_t NX = 0, NY = 0, Pitch = 0, Npad = 0, Ndense = 0, NMAX = 0, nshots = 0;
    std::vector<int64_t> order;         // plan position -> caller shot
    std::vector<int64_t> nstep_sorted;  // per plan position
    std::vector<int64_t> srcv_off_caller, trace_off_caller;  // per caller shot
    int64_t srcv_elems = 0, trace_elems = 0;
    std::vector<int> active;  // active[n] = #shots with nstep >= n
    std::vector<double> sx, ty;
    // phi/psi vanish on the quiet box (no damping profile acts there): checkpoint records keep
    // the nfr cells outside it
    FrameBox fbox{};
    int64_t nfr = 0;
    // snapshot slots of the reversal schedule (< 0: one per segment start, the one-level scheme)
    int64_t snapshots = -1;
    std::vector<uint8_t> pmlx, pmly, nearx, neary;
    double f_pml = 0.0;  // fraction of interior cells inside the PML strips
    // defaults: the benchmark shape (forward: the streaming two-level rectangle)
    Shape fwd_shape{8, 16, 1}, adj_shape{2, 16, 1};
    int tuned = 0;
    // device tables; cell tables exist for the pitched (plan) and dense (history) layouts
    DevBuf d_sx, d_ty, d_pmlx, d_pmly, d_nearx, d_neary;
    DevBuf d_us_ptr, d_us_eptr, d_us_ent, d_nsrc, d_srcv_off;
    DevBuf d_rcv_ptr, d_trace_off;
    DevBuf d_us_cell[2], d_rcv_cell[2];  // [0] pitched, [1] dense
    // receiver injection tables for the current adjoint shot grouping (inj_sg)
    std::vector<int64_t> rcv_i0, rcv_j0, rcv_ptr_h;  // 0-based receiver (i, j), plan order
    // per union cell: segments [gi_sptr[c], gi_sptr[c+1]) of (shot gi_sshot, entries
    // [gi_seg[q], gi_seg[q+1]) into gi_rcv)
    DevBuf d_gi_ptr, d_gi_sptr, d_gi_sshot, d_gi_seg, d_gi_rcv, d_gi_cell[2];
    int inj_sg = 0;
    int64_t max_gi_cells = 0;
    // interior rectangle and frame work items, per layout ([0] pitched, [1] dense)
    Rect rect[2]{};
    DevBuf d_items[2];
    int nitems[2] = {0, 0};
    // Two-level forward (temporal blocking) in the pitched layout: rect1 is rect[0] shrunk by
    // one cell and items1 the frame items of its complement (level n+1's frame launch of a
    // kernel-0 pair).
    // blk_ok: rect1 is not empty and no source lies in rect[0]; blocking: levels per
    // rectangle pass asked for (1 or 2).
    Rect rect1{};
    DevBuf d_items1;
    int nitems1 = 0;
    bool blk_ok = false;
    int blocking = 2;
    // sources added inside the frame work (kernel 1, no source in rect[0]); receivers then
    // recorded per forward segment (ac_record)
    bool src_fused = false;
    // items of the two-level frame pass (complement of rect1, WAVE - 2 columns each)
    DevBuf d_items2[FRAME2_H];  // two-level frame items of height frame2_height(k)
    int nitems2[FRAME2_H] = {};
    int frame2_rows = 0;        // forced item height (ads_acoustic_plan_set_frame_rows), 0: chosen per launch
    // kernel 0 with the sources in the frame work: items of WAVE - 2 columns covering the
    // complement of rect[0] and of rect1, for the frame body that also updates phi/psi
    DevBuf d_items_h, d_items1h;
    int nitems_h = 0, nitems1h = 0;
    DevBuf d_sc_idx, d_sc_ptr, d_sc_shot, d_sc_base, d_sc_ns, d_nstep_sorted;
    // hipGraph replay of whole checkpoint segments (forward sweep, recomputation, adjoint
    // sweep), keyed by everything the captured launches depend on
    struct GraphKey {
        int kind, record;
        int64_t lo, hi, ns;  // ns: slots of the workspace layout (fixes every slot address)
        const void *ws, *srcv, *traces, *resid;
        bool operator<(const GraphKey &o) const
        {
            return std::tie(kind, record, lo, hi, ns, ws, srcv, traces, resid) <
                   std::tie(o.kind, o.record, o.lo, o.hi, o.ns, o.ws, o.srcv, o.traces, o.resid);
        }
    };
    // A cached graph also records its last use: the key holds the caller's trace / residual
    // addresses, so fresh buffers on every FWI iteration would otherwise add graphs without
    // bound.  The cache keeps at most graph_cap() graphs and evicts the least recently used.
    struct CachedGraph {
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::map<GraphKey, CachedGraph> graphs;
    uint64_t graph_clock = 0;
    hipStream_t cap_stream = nullptr;  // capture needs a stream other than the null stream
    int use_graph = 0;
    void drop_graphs()
    {
        if (!graphs.empty()) (void)hipDeviceSynchronize();  // none may still be executing
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second.exec);
        graphs.clear();
    }
    // one gradient's worth of segment graphs (forward, recomputation and adjoint sweeps of M
    // segments) for two buffer sets, so alternating calls keep their graphs
    static size_t graph_cap(int64_t M) { return (size_t)std::max<int64_t>(16, 6 * M + 8); }
    void evict_lru()
    {
        auto victim = graphs.begin();
        for (auto it = graphs.begin(); it != graphs.end(); ++it)
            if (it->second.used < victim->second.used) victim = it;
        (void)hipDeviceSynchronize();  // the evicted graph may still be executing
        (void)hipGraphExecDestroy(victim->second.exec);
        graphs.erase(victim);
    }
    virtual ~AcPlanBase()
    {
        drop_graphs();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

static int validate_params(const ads_acoustic_params *p)
{
    if (!p) ADS_FAIL(ADS_ERR_ARG, "null params");
    if (p->NX < 1 || p->NY < 1 || p->NSTEP < 1)
        ADS_FAIL(ADS_ERR_SHAPE, "NX, NY, NSTEP must be >= 1 (got %lld, %lld, %lld)",
                 (long long)p->NX, (long long)p->NY, (long long)p->NSTEP);
    if ((p->NX + 2) * (((p->NY + 2 + 63) / 64) * 64) >= (int64_t)1 << 31 || p->NX + 2 > (1 << 24))
        ADS_FAIL(ADS_ERR_SHAPE, "grid too large for 32-bit cell indexing within a shot");
    if (!(p->DELTAT > 0) || !(p->DELTAX > 0) || !(p->DELTAY > 0))
        ADS_FAIL(ADS_ERR_ARG, "DELTAT, DELTAX, DELTAY must be > 0");
    if (p->PropagatorKernel < 0 || p->PropagatorKernel > 2)
        ADS_FAIL(ADS_ERR_ARG, "PropagatorKernel must be 0, 1 or 2");
    return ADS_OK;
}

template <typename T> struct AcPlan : AcPlanBase {
    AcCoef<T> k;

    struct Layout {  // pitched (plan) or dense (history) addressing
        int P;
        int li;  // index into the cell tables
    };
    Layout plan_layout() const { return Layout{(int)Pitch, 0}; }
    Layout dense_layout() const { return Layout{(int)(NY + 2), 1}; }

    static unsigned grid_of(const TileMap &m) { return cdiv(m.nb, NXCD) * NXCD; }

    static constexpr int VECP = sizeof(T) == 4 ? 4 : 2;  // pitched layout: 16-byte vectors
    static TileMap rect_tiles(const Rect &R, int vec, int rpt, int act, int sg)
    {
        TileMap m;
        m.ntj = (int)cdiv(R.jR - R.jL, WAVE * vec);
        m.nstrips = (int)cdiv(R.iR - R.iL, NWAVES * rpt);
        m.ngroups = (int)cdiv(act, sg);
        m.nb = m.ntj * m.nstrips * m.ngroups;
        return m;
    }
    static TileMap item_tiles(int n, int act, int sg)
    {
        TileMap m;
        m.ntj = 1;
        m.nstrips = n;
        m.ngroups = (int)cdiv(act, sg);
        m.nb = m.ntj * m.nstrips * m.ngroups;
        return m;
    }
    // kernel 0 in the pitched layout with the sources from the per-cell table: its frame work
    // also updates phi/psi and adds the sources (fwd_frame0h_body)
    bool frame_h(const FwdArgs<T> &a, int li) const
    {
        return variant == 0 && li == 0 && a.src.cell_idx != nullptr && nitems_h > 0;
    }
    TileMap frame_tiles(int li, int act, int sg) const
    {
        TileMap m;
        m.ntj = 1;
        m.nstrips = nitems[li];
        m.ngroups = (int)cdiv(act, sg);
        m.nb = m.ntj * m.nstrips * m.ngroups;
        return m;
    }
    static bool empty(const Rect &R) { return R.iR <= R.iL || R.jR <= R.jL; }
    // rectangle kernel instances: rows per wavefront 1/2/4, shots per batch 1/2 (2 with 1-2 rows).
    // With `fuse` (pitched layout only) the frame items of layout li ride in the same launch.
    template <int RPT, int VEC, int SB>
    void launch_fwd_rect(FwdArgs<T> a, const Rect &R, int li, bool fuse, hipStream_t st)
    {
        a.map = rect_tiles(R, VEC, RPT, a.act, a.Sg);
        if constexpr (VEC > 1) {
            if (fuse) {
                const bool h = frame_h(a, li);
                const int ni = h ? nitems_h : nitems[li];
                const TileMap fm = item_tiles(ni, a.act, a.Sg);
                const int nfb = (int)grid_of(fm);
                const unsigned nb = (unsigned)nfb + grid_of(a.map);
                const int4 *it = h ? d_items_h.template as<int4>() : d_items[li].template as<int4>();
                if (variant == 1)
                    ac_fwd_level<T, RPT, VEC, SB, 1><<<nb, WAVE * NWAVES, 0, st>>>(a, R, fm, it, ni, nfb);
                else if (h)
                    ac_fwd_level<T, RPT, VEC, SB, 2><<<nb, WAVE * NWAVES, 0, st>>>(a, R, fm, it, ni, nfb);
                else
                    ac_fwd_level<T, RPT, VEC, SB, 0><<<nb, WAVE * NWAVES, 0, st>>>(a, R, fm, it, ni, nfb);
                return;
            }
        }
        ac_fwd_rect<T, RPT, VEC, SB><<<grid_of(a.map), WAVE * NWAVES, 0, st>>>(a, R);
    }
    template <int VEC>
    void launch_fwd_shape(const FwdArgs<T> &a, const Rect &R, const Shape &sh, int li, bool fuse, hipStream_t st)
    {
        if (empty(R)) return;
        if (sh.sb == 2 && sh.rpt == 1) launch_fwd_rect<1, VEC, 2>(a, R, li, fuse, st);
        else if (sh.sb == 2 && sh.rpt == 2) launch_fwd_rect<2, VEC, 2>(a, R, li, fuse, st);
        else if (sh.rpt == 1) launch_fwd_rect<1, VEC, 1>(a, R, li, fuse, st);
        else if (sh.rpt == 4) launch_fwd_rect<4, VEC, 1>(a, R, li, fuse, st);
        else launch_fwd_rect<2, VEC, 1>(a, R, li, fuse, st);
    }
    template <int RPT, int VEC, int SB>
    void launch_adj_rect(AdjArgs<T> a, const Rect &R, int li, bool fuse, hipStream_t st)
    {
        a.map = rect_tiles(R, VEC, RPT, a.act, a.Sg);
        if constexpr (VEC > 1) {
            if (fuse) {
                const TileMap fm = frame_tiles(li, a.act, a.Sg);
                const int nfb = (int)grid_of(fm);
                const unsigned nb = (unsigned)nfb + grid_of(a.map);
                const int4 *it = d_items[li].template as<int4>();
                if (variant == 1)
                    ac_adj_level<T, RPT, VEC, SB, 1><<<nb, WAVE * NWAVES, 0, st>>>(a, R, fm, it, nitems[li], nfb);
                else
                    ac_adj_level<T, RPT, VEC, SB, 0><<<nb, WAVE * NWAVES, 0, st>>>(a, R, fm, it, nitems[li], nfb);
                return;
            }
        }
        ac_adj_rect<T, RPT, VEC, SB><<<grid_of(a.map), WAVE * NWAVES, 0, st>>>(a, R);
    }
    template <int VEC>
    void launch_adj_shape(const AdjArgs<T> &a, const Rect &R, const Shape &sh, int li, bool fuse, hipStream_t st)
    {
        if (empty(R)) return;
        if (sh.sb == 2 && sh.rpt == 1) launch_adj_rect<1, VEC, 2>(a, R, li, fuse, st);
        else if (sh.sb == 2 && sh.rpt == 2) launch_adj_rect<2, VEC, 2>(a, R, li, fuse, st);
        else if (sh.rpt == 1) launch_adj_rect<1, VEC, 1>(a, R, li, fuse, st);
        else if (sh.rpt == 4) launch_adj_rect<4, VEC, 1>(a, R, li, fuse, st);
        else launch_adj_rect<2, VEC, 1>(a, R, li, fuse, st);
    }
    // The rectangle and the frame (disjoint cells, same inputs): one fused launch in the
    // pitched layout, two launches in the dense history layout.
    void launch_fwd_all(FwdArgs<T> a, const Layout &L, const Shape &sh, hipStream_t st)
    {
        const bool fuse = L.li == 0 && !empty(rect[0]) && nitems[0] > 0;
        if (L.li == 0) launch_fwd_shape<VECP>(a, rect[0], sh, 0, fuse, st);
        else launch_fwd_shape<1>(a, rect[1], sh, 1, false, st);
        if (fuse || nitems[L.li] == 0) return;
        const bool h = frame_h(a, L.li);
        const int ni = h ? nitems_h : nitems[L.li];
        a.map = item_tiles(ni, a.act, a.Sg);
        const int4 *it = h ? d_items_h.template as<int4>() : d_items[L.li].template as<int4>();
        if (variant == 1) ac_fwd_frame<T, 1><<<grid_of(a.map), WAVE * NWAVES, 0, st>>>(a, it, ni);
        else if (h) ac_fwd_frame<T, 2><<<grid_of(a.map), WAVE * NWAVES, 0, st>>>(a, it, ni);
        else ac_fwd_frame<T, 0><<<grid_of(a.map), WAVE * NWAVES, 0, st>>>(a, it, ni);
    }
    void launch_adj_all(AdjArgs<T> a, const Layout &L, const Shape &sh, hipStream_t st)
    {
        const bool fuse = L.li == 0 && !empty(rect[0]) && nitems[0] > 0;
        if (L.li == 0) launch_adj_shape<VECP>(a, rect[0], sh, 0, fuse, st);
        else launch_adj_shape<1>(a, rect[1], sh, 1, false, st);
        if (fuse || nitems[L.li] == 0) return;
        a.map = frame_tiles(L.li, a.act, a.Sg);
        const int4 *it = d_items[L.li].template as<int4>();
        if (variant == 1) ac_adj_frame<T, 1><<<grid_of(a.map), WAVE * NWAVES, 0, st>>>(a, it, nitems[L.li]);
        else ac_adj_frame<T, 0><<<grid_of(a.map), WAVE * NWAVES, 0, st>>>(a, it, nitems[L.li]);
    }

    double fwd_level_bytes(int act) const
    {
        return (double)act * (double)NX * (double)NY * sizeof(T) * (4.0 + (variant == 1 ? 4.0 : 2.0) * f_pml);
    }
    // item list of the two-level frame work for a launch over act shots in groups of Sg
    int frame2_index(int act, int Sg) const
    {
        for (int k = 0; k < FRAME2_H; k++)
            if (frame2_height(k) == frame2_rows) return k;
        return frame2_pick(act, (int)cdiv(std::min(act, Sg), NWAVES), 2.0 * fwd_level_bytes(act));
    }
    FwdArgs<T> fwd_args(const Layout &L, int act, Lv<T> uo, Lv<const T> w, Lv<const T> wo,
                        Lv<const T> phi, Lv<const T> psi, Lv<T> phio, Lv<T> psio, const T *C,
                        const Shape &shape, int n = 0, const T *srcv = nullptr) const
    {
        SrcIn<T> src{};
        if (src_fused && L.li == 0 && srcv)
            src = SrcIn<T>{d_sc_idx.as<int32_t>(), d_sc_ptr.as<int32_t>(), d_sc_shot.as<int32_t>(),
                           d_sc_base.as<int64_t>(), d_sc_ns.as<int64_t>(), srcv, n};
        return FwdArgs<T>{uo, w, wo, phi, psi, phio, psio, C, d_sx.as<T>(), d_ty.as<T>(),
                          d_nearx.as<uint8_t>(), d_neary.as<uint8_t>(), (int)NX, (int)NY, L.P, act,
                          std::min(shape.sg, act), TileMap{}, k, src};
    }
    // after a level's stencil launch: phi/psi from the new u (kernel 0), sources, receivers
    void fwd_post(const Layout &L, const FwdArgs<T> &a, int n, const T *srcv, T *traces, int record,
                  hipStream_t st)
    {
        if (a.src.cell_idx) return;  // sources added by the frame work, receivers per segment
        if (variant == 0) {
            dim3 g((unsigned)nitems[L.li], cdiv(a.act, NWAVES));
            ac_fwd_phipsi<T><<<g, 256, 0, st>>>(cst(a.u), a.phi, a.psi, a.phio, a.psio, a.C, d_sx.as<T>(),
                                                d_ty.as<T>(), d_pmlx.as<uint8_t>(),
                                                d_pmly.as<uint8_t>(), (int)NX, (int)NY, L.P, k,
                                                d_items[L.li].template as<int4>(), a.act);
        }
        ac_src_rcv<T><<<a.act, 256, 0, st>>>(a.u, n, d_us_ptr.as<int64_t>(), d_us_cell[L.li].template as<int64_t>(),
                                             d_us_eptr.as<int64_t>(), d_us_ent.as<int64_t>(), srcv,
                                             d_srcv_off.as<int64_t>(), d_nsrc.as<int64_t>(), k.dt2,
                                             d_rcv_ptr.as<int64_t>(), d_rcv_cell[L.li].template as<int64_t>(),
                                             traces, d_trace_off.as<int64_t>(), record);
    }

    int fwd_step(const Layout &L, int n, int act, Lv<T> uo, Lv<const T> w, Lv<const T> wo,
                 Lv<const T> phi, Lv<const T> psi, Lv<T> phio, Lv<T> psio, const T *C,
                 const T *srcv, T *traces, int record, hipStream_t st, const Shape *sh = nullptr)
    {
        if (act <= 0) return ADS_OK;
        const Shape &shape = sh ? *sh : fwd_shape;
        FwdArgs<T> a = fwd_args(L, act, uo, w, wo, phi, psi, phio, psio, C, shape, n, srcv);
        int pr = prof_start(ADS_PROF_AC_FWD, st);
        launch_fwd_all(a, L, shape, st);
        prof_stop(pr, st, fwd_level_bytes(act));
        fwd_post(L, a, n, srcv, traces, record, st);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int adj_level(const Layout &L, int n, int act, Lv<T> P, Lv<const T> P1, Lv<T> F0, Lv<T> G0,
                  Lv<const T> F1, Lv<const T> G1, Lv<const T> um1, Lv<const T> un,
                  Lv<const T> dense, Lv<T> gC, const T *C, const T *srcv, const T *resid,
                  hipStream_t st, const Shape *sh = nullptr)
    {
        if (act <= 0) return ADS_OK;
        const Shape &shape = sh ? *sh : adj_shape;
        const int Sg = adj_shape.sg;  // gC grouping is fixed per plan (see group_arrays())
        const double cells = (double)act * (double)NX * (double)NY;
        if (variant == 0) {
            dim3 g((unsigned)nitems[L.li], cdiv(act, std::min(Sg, act)));
            ac_adj_phipsi<T><<<g, 256, 0, st>>>(
                P1, F1, G1, F0, G0, un, gC, std::min(Sg, act), act, n, d_sx.as<T>(), d_ty.as<T>(), d_pmlx.as<uint8_t>(),
                d_pmly.as<uint8_t>(), d_us_ptr.as<int64_t>(), d_us_cell[L.li].template as<int64_t>(),
                d_us_eptr.as<int64_t>(), d_us_ent.as<int64_t>(), srcv, d_srcv_off.as<int64_t>(),
                d_nsrc.as<int64_t>(), (int)NX, (int)NY, L.P, k, d_items[L.li].template as<int4>());
        }
        AdjArgs<T> a{P, P1, variant == 1 ? F1 : cst(F0), variant == 1 ? G1 : cst(G0), F0, G0,
                     um1, dense, gC, C, d_sx.as<T>(), d_ty.as<T>(), d_nearx.as<uint8_t>(),
                     d_neary.as<uint8_t>(), (int)NX, (int)NY, L.P, act, std::min(Sg, act),
                     TileMap{}, k};
        int pr = prof_start(ADS_PROF_AC_ADJ, st);
        launch_adj_all(a, L, shape, st);
        prof_stop(pr, st, cells * sizeof(T) * (7.0 + (variant == 1 ? 4.0 : 2.0) * f_pml));
        if (resid && max_gi_cells > 0) {
            if (inj_sg != Sg) ADS_TRY(build_inject_tables(Sg));
            dim3 g(cdiv(max_gi_cells, NWAVES), cdiv(act, std::min(Sg, act)));
            ac_adj_inject<T><<<g, WAVE * NWAVES, 0, st>>>(P, um1, gC, act, n, d_gi_ptr.as<int64_t>(),
                                                d_gi_cell[L.li].template as<int64_t>(),
                                                d_gi_sptr.as<int64_t>(), d_gi_sshot.as<int32_t>(),
                                                d_gi_seg.as<int64_t>(),
                                                d_gi_rcv.as<int64_t>(), d_rcv_ptr.as<int64_t>(),
                                                resid, d_trace_off.as<int64_t>(), d_sx.as<T>(),
                                                d_ty.as<T>(), (int)NX, (int)NY, L.P, k);
        }
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int64_t group_arrays() const { return (int64_t)cdiv(nshots, adj_shape.sg); }

    // Union of receiver cells per shot group of size Sg (plan order); per cell, entries sorted
    // by (shot, receiver order).  Note: the Sg used at a level with act < Sg shots is act,
    // which puts every active shot in group 0 -- the same as these tables (shots < Sg).
    int build_inject_tables(int Sg)
    {
        if (Sg < 1) Sg = 1;
        const int64_t ng = (int64_t)cdiv(nshots, Sg);
        std::vector<int64_t> gptr(1, 0), sptr(1, 0), seg(1, 0), rcv, cells[2];
        std::vector<int32_t> sshot;
        max_gi_cells = 0;
        for (int64_t g = 0; g < ng; g++) {
            struct E { int64_t i, j; int32_t s; int64_t r; };
            std::vector<E> es;
            for (int64_t s = g * Sg; s < std::min<int64_t>(nshots, (g + 1) * Sg); s++)
                for (int64_t e = rcv_ptr_h[s]; e < rcv_ptr_h[s + 1]; e++)
                    es.push_back(E{rcv_i0[e], rcv_j0[e], (int32_t)s, e - rcv_ptr_h[s]});
            std::stable_sort(es.begin(), es.end(), [](const E &a, const E &b) {
                if (a.i != b.i) return a.i < b.i;
                if (a.j != b.j) return a.j < b.j;
                return a.s < b.s;  // stable: receiver order kept within a shot
            });
            int64_t ncell = 0;
            for (size_t q = 0; q < es.size();) {
                size_t z = q;
                while (z < es.size() && es[z].i == es[q].i && es[z].j == es[q].j) {
                    if (z == q || es[z].s != es[z - 1].s) {  // a new shot segment
                        if (z != q) seg.push_back((int64_t)rcv.size());
                        sshot.push_back(es[z].s);
                    }
                    rcv.push_back(es[z].r);
                    z++;
                }
                seg.push_back((int64_t)rcv.size());
                sptr.push_back((int64_t)sshot.size());
                cells[0].push_back(es[q].i * Pitch + es[q].j);
                cells[1].push_back(es[q].i * (NY + 2) + es[q].j);
                ncell++;
                q = z;
            }
            gptr.push_back(gptr.back() + ncell);
            max_gi_cells = std::max(max_gi_cells, ncell);
        }
        ADS_TRY(d_gi_ptr.upload(gptr));
        ADS_TRY(d_gi_sptr.upload(sptr));
        ADS_TRY(d_gi_sshot.upload(sshot));
        ADS_TRY(d_gi_seg.upload(seg));
        ADS_TRY(d_gi_rcv.upload(rcv));
        ADS_TRY(d_gi_cell[0].upload(cells[0]));
        ADS_TRY(d_gi_cell[1].upload(cells[1]));
        inj_sg = Sg;
        return ADS_OK;
    }

    // ---- workspace layout for checkpoint interval K ----
    struct WS {
        int64_t K, NS, M, S;  // interval, slots, segments, snapshot slots
        // phi/psi of level n in buffer n mod 3: a two-level pass writes level n+1 while other
        // workgroups of the same launch still read level n-1
        T *seg, *phi[3], *psi[3], *P[2], *F[2], *G[2], *gC, *ckpt, *C, *resid;
        double *loss, *part;
        std::vector<size_t> ck_off;  // element offset of each snapshot slot in ckpt
    };
    int64_t seg_level(const WS &w, int64_t m) const { return std::min<int64_t>(1 + m * w.K, NMAX); }
    int64_t snap_slots(int64_t M, int64_t snaps) const
    {
        const int64_t one = std::max<int64_t>(M - 1, 0);
        return snaps < 0 ? one : std::min(snaps, one);
    }
    // a record of the state at level a for `act` shots: u^a, u^{a-1} (full grid), phi^a and
    // psi^a (cells outside the quiet box)
    size_t rec_elems(int act) const { return (size_t)act * (2 * (size_t)Npad + 2 * (size_t)nfr); }
    int rec_act(int64_t a) const { return active[std::min<int64_t>(a + 1, NMAX)]; }
    int64_t seg_count(int64_t K) const { return NMAX >= 2 ? (NMAX - 1 + K - 1) / K : 0; }
    void layout(WS &w, Carve &cv, int64_t K) const { layout(w, cv, K, snapshots); }
    void layout(WS &w, Carve &cv, int64_t K, int64_t snaps) const
    {
        if (K < 1) K = 1;
        if (K > std::max<int64_t>(NMAX - 1, 1)) K = std::max<int64_t>(NMAX - 1, 1);
        w.K = K;
        w.M = seg_count(K);
        w.NS = std::min<int64_t>(K + 2, NMAX + 1);
        const size_t F = (size_t)nshots * Npad * sizeof(T);
        w.seg = (T *)cv.take(F * w.NS);
        for (int q = 0; q < 3; q++) {
            w.phi[q] = (T *)cv.take(F);
            w.psi[q] = (T *)cv.take(F);
        }
        for (int q = 0; q < 2; q++) {
            w.P[q] = (T *)cv.take(F);
            w.F[q] = (T *)cv.take(F);
            w.G[q] = (T *)cv.take(F);
        }
        w.gC = (T *)cv.take((size_t)nshots * Npad * sizeof(T));  // up to one array per shot
        // snapshot slots: one-level, slot q holds state q+1 and is sized for the shots active
        // there; binomial, any slot may hold any state and is sized for the most active shots
        w.S = snap_slots(w.M, snaps);
        w.ck_off.assign((size_t)w.S, 0);
        size_t ck = 0;
        for (int64_t q = 0; q < w.S; q++) {
            w.ck_off[q] = ck;
            const int64_t a = w.S >= w.M - 1 ? 1 + (q + 1) * w.K : 1 + w.K;
            ck += (rec_elems(rec_act(a)) + 63) & ~size_t(63);
        }
        w.ckpt = (T *)cv.take(ck * sizeof(T));
        w.C = (T *)cv.take((size_t)Npad * sizeof(T));
        w.resid = (T *)cv.take((size_t)std::max<int64_t>(trace_elems, 1) * sizeof(T));
        w.loss = (double *)cv.take(64);
        w.part = (double *)cv.take(MISFIT_BLOCKS * sizeof(double));
    }
    Lv<T> phl(const WS &w, int64_t n) const { return Lv<T>{w.phi[n % 3], Npad}; }
    Lv<T> psl(const WS &w, int64_t n) const { return Lv<T>{w.psi[n % 3], Npad}; }
    size_t ws_size(int64_t K) const { return ws_size(K, snapshots); }
    size_t ws_size(int64_t K, int64_t snaps) const
    {
        WS w;
        Carve cv(nullptr, 0);
        layout(w, cv, K, snaps);
        return cv.off + 256;
    }

    // Fastest (K, S) whose workspace fits `budget`.  Cost model in shot-levels: the forward
    // and the adjoint sweep each pass every active shot's levels once (an adjoint level costs
    // ADJ_COST forward levels), and the reversal schedule adds its segment rebuilds.
    static constexpr double ADJ_COST = 1.7;
    int choose_schedule(size_t budget, int64_t *Kout, int64_t *Sout) const
    {
        double total = 0.0;
        for (int64_t n = 2; n <= NMAX; n++) total += active[n];
        auto ws = [&](int64_t K, int64_t S) { return ws_size(K, S); };
        auto work = [&](int64_t K, int64_t M) {
            std::vector<double> w((size_t)M, 0.0);
            for (int64_t m = 0; m < M; m++) {
                const int64_t a = std::min<int64_t>(1 + m * K, NMAX), b = std::min<int64_t>(a + K, NMAX);
                for (int64_t n = a + 1; n <= b; n++) w[m] += active[n];
            }
            return w;
        };
        return ads::choose_schedule(budget, NMAX - 1, total, ADJ_COST, ws, work, *Kout, *Sout)
                   ? ADS_OK
                   : ADS_ERR_WORKSPACE;
    }

    Lv<T> slot(const WS &w, int64_t n) const
    {
        return Lv<T>{w.seg + (size_t)(n % w.NS) * nshots * Npad, Npad};
    }

    int prepare_C(T *C, const void *model, int P, hipStream_t st)
    {
        k_model_in<T><<<dim3(cdiv(NY + 2, 256), NX + 2), 256, 0, st>>>((const T *)model, C, (int)NX,
                                                                      (int)NY, P, p.square_velocity);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    // forward over levels lo+1..hi from the state held in slots/phi (record traces if asked)
    // Runs body(stream) directly, or -- with graphs on and no event profiling -- captures it
    // once on a private stream into a graph and replays that graph on st.
    template <class F> int replay(const GraphKey &key, int64_t M, hipStream_t st, F &&body)
    {
        if (!use_graph || prof_enabled()) return body(st);
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            while (!graphs.empty() && graphs.size() >= graph_cap(M)) evict_lru();
            if (!cap_stream) ADS_HIP(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
            ADS_HIP(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
            const int rc = body(cap_stream);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(cap_stream, &g);
            if (rc != ADS_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            ADS_HIP(e);
            hipGraphExec_t ex = nullptr;
            const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            ADS_HIP(ei);
            it = graphs.emplace(key, CachedGraph{ex, 0}).first;
        }
        it->second.used = ++graph_clock;
        ADS_HIP(hipGraphLaunch(it->second.exec, st));
        return ADS_OK;
    }

    // Levels n and n+1 of the plan layout with the two-level rectangle pass.  Kernel 1 with the
    // sources added by the frame work: one launch (two-level rectangle + two-level frame).
    // Otherwise: [rectangle levels n, n+1 + frame level n], phi/psi and sources of level n,
    // [frame of the shrunk rectangle, level n+1], phi/psi and sources of level n+1.
    template <int ROWS>
    void launch_level2_s(FwdArgs<T> a, Lv<T> u2, hipStream_t st)
    {
        const Rect &R = rect[0];
        a.map.ntj = (int)cdiv(R.jR - R.jL, WAVE * VECP);
        a.map.nstrips = (int)cdiv(R.iR - R.iL, ROWS);
        a.map.ngroups = (int)cdiv(a.act, NWAVES);
        a.map.nb = a.map.ntj * a.map.nstrips * a.map.ngroups;
        const bool h = frame_h(a, 0);
        const int ni = h ? nitems_h : nitems[0];
        const TileMap fm = item_tiles(ni, a.act, a.Sg);
        const int nfb = (int)grid_of(fm);
        const unsigned nb = (unsigned)nfb + grid_of(a.map);
        const int4 *it = h ? d_items_h.template as<int4>() : d_items[0].template as<int4>();
        if (variant == 1)
            ac_fwd_level2_s<T, VECP, ROWS, 1><<<nb, WAVE * NWAVES, 0, st>>>(a, u2, R, fm, it, ni, nfb);
        else if (h)
            ac_fwd_level2_s<T, VECP, ROWS, 2><<<nb, WAVE * NWAVES, 0, st>>>(a, u2, R, fm, it, ni, nfb);
        else
            ac_fwd_level2_s<T, VECP, ROWS, 0><<<nb, WAVE * NWAVES, 0, st>>>(a, u2, R, fm, it, ni, nfb);
    }
    template <int RPT>
    void launch_level2(FwdArgs<T> a, Lv<T> u2, hipStream_t st)
    {
        const Rect &R = rect[0];
        a.map = rect_tiles(R, VECP, RPT, a.act, a.Sg);
        const bool h = frame_h(a, 0);
        const int ni = h ? nitems_h : nitems[0];
        const TileMap fm = item_tiles(ni, a.act, a.Sg);
        const int nfb = (int)grid_of(fm);
        const unsigned nb = (unsigned)nfb + grid_of(a.map);
        const int4 *it = h ? d_items_h.template as<int4>() : d_items[0].template as<int4>();
        if (variant == 1)
            ac_fwd_level2<T, RPT, VECP, 1><<<nb, WAVE * NWAVES, 0, st>>>(a, u2, R, fm, it, ni, nfb);
        else if (h)
            ac_fwd_level2<T, RPT, VECP, 2><<<nb, WAVE * NWAVES, 0, st>>>(a, u2, R, fm, it, ni, nfb);
        else
            ac_fwd_level2<T, RPT, VECP, 0><<<nb, WAVE * NWAVES, 0, st>>>(a, u2, R, fm, it, ni, nfb);
    }
    // Levels n and n+1 of the whole grid in one launch (kernel 1 with the frame adding sources)
    template <int RPT>
    void launch_pair(FwdArgs<T> a, Lv<T> u2, Lv<T> phi2, Lv<T> psi2, hipStream_t st)
    {
        const Rect &R = rect[0];
        a.map = rect_tiles(R, VECP, RPT, a.act, a.Sg);
        TileMap fm;
        fm.ntj = 1;
        const int h = frame2_index(a.act, a.Sg);
        fm.nstrips = nitems2[h];
        fm.ngroups = (int)cdiv(a.act, a.Sg);
        fm.nb = fm.ntj * fm.nstrips * fm.ngroups;
        const int nfb = (int)grid_of(fm);
        const unsigned nb = (unsigned)nfb + grid_of(a.map);
        ac_fwd_pair<T, RPT, VECP><<<nb, WAVE * NWAVES, 0, st>>>(a, u2, phi2, psi2, R, fm,
                                                                d_items2[h].template as<int4>(), nitems2[h], nfb);
    }
    template <int ROWS>
    void launch_pair_s(FwdArgs<T> a, Lv<T> u2, Lv<T> phi2, Lv<T> psi2, hipStream_t st)
    {
        const Rect &R = rect[0];
        a.map.ntj = (int)cdiv(R.jR - R.jL, WAVE * VECP);
        a.map.nstrips = (int)cdiv(R.iR - R.iL, ROWS);
        a.map.ngroups = (int)cdiv(a.act, NWAVES);
        a.map.nb = a.map.ntj * a.map.nstrips * a.map.ngroups;
        TileMap fm;
        fm.ntj = 1;
        const int h = frame2_index(a.act, a.Sg);
        fm.nstrips = nitems2[h];
        fm.ngroups = (int)cdiv(a.act, a.Sg);
        fm.nb = fm.ntj * fm.nstrips * fm.ngroups;
        const int nfb = (int)grid_of(fm);
        const unsigned nb = (unsigned)nfb + grid_of(a.map);
        ac_fwd_pair_s<T, VECP, ROWS><<<nb, WAVE * NWAVES, 0, st>>>(a, u2, phi2, psi2, R, fm,
                                                                   d_items2[h].template as<int4>(), nitems2[h], nfb);
    }
    int fwd_pair(const WS &w, int n, int act, const T *srcv, T *traces, int record, hipStream_t st,
                 const Shape *sh)
    {
        const Layout L = plan_layout();
        const Shape &shape = sh ? *sh : fwd_shape;
        FwdArgs<T> a = fwd_args(L, act, slot(w, n), cst(slot(w, n - 1)), cst(slot(w, n - 2)),
                                cst(phl(w, n - 1)), cst(psl(w, n - 1)), phl(w, n), psl(w, n), w.C, shape,
                                n, srcv);
        if (variant == 1 && nitems2[0] > 0 && a.src.cell_idx) {
            int pr = prof_start(ADS_PROF_AC_FWD, st);
            if (shape.rpt == 1) launch_pair<1>(a, slot(w, n + 1), phl(w, n + 1), psl(w, n + 1), st);
            else if (shape.rpt == 4) launch_pair<4>(a, slot(w, n + 1), phl(w, n + 1), psl(w, n + 1), st);
            else if (shape.rpt == 8) launch_pair_s<8>(a, slot(w, n + 1), phl(w, n + 1), psl(w, n + 1), st);
            else if (shape.rpt == 16) launch_pair_s<16>(a, slot(w, n + 1), phl(w, n + 1), psl(w, n + 1), st);
            else launch_pair<2>(a, slot(w, n + 1), phl(w, n + 1), psl(w, n + 1), st);
            prof_stop(pr, st, 2.0 * fwd_level_bytes(act), 2);
            ADS_LAUNCH_CHECK();
            return ADS_OK;
        }
        int pr = prof_start(ADS_PROF_AC_FWD, st);
        if (shape.rpt == 1) launch_level2<1>(a, slot(w, n + 1), st);
        else if (shape.rpt == 4) launch_level2<4>(a, slot(w, n + 1), st);
        else if (shape.rpt == 8) launch_level2_s<8>(a, slot(w, n + 1), st);
        else if (shape.rpt == 16) launch_level2_s<16>(a, slot(w, n + 1), st);
        else launch_level2<2>(a, slot(w, n + 1), st);
        prof_stop(pr, st, fwd_level_bytes(act));
        fwd_post(L, a, n, srcv, traces, record, st);
        FwdArgs<T> b = fwd_args(L, act, slot(w, n + 1), cst(slot(w, n)), cst(slot(w, n - 1)),
                                cst(phl(w, n)), cst(psl(w, n)), phl(w, n + 1), psl(w, n + 1), w.C, shape,
                                n + 1, srcv);
        pr = prof_start(ADS_PROF_AC_FWD, st);
        if (nitems1 > 0) {
            const bool h = frame_h(b, 0);
            const int ni = h ? nitems1h : nitems1;
            b.map = item_tiles(ni, act, b.Sg);
            const int4 *it = h ? d_items1h.template as<int4>() : d_items1.template as<int4>();
            if (variant == 1) ac_fwd_frame<T, 1><<<grid_of(b.map), WAVE * NWAVES, 0, st>>>(b, it, ni);
            else if (h) ac_fwd_frame<T, 2><<<grid_of(b.map), WAVE * NWAVES, 0, st>>>(b, it, ni);
            else ac_fwd_frame<T, 0><<<grid_of(b.map), WAVE * NWAVES, 0, st>>>(b, it, ni);
        }
        prof_stop(pr, st, fwd_level_bytes(act));
        fwd_post(L, b, n + 1, srcv, traces, record, st);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int run_forward(const WS &w, int64_t lo, int64_t hi, const T *srcv, T *traces, int record,
                    hipStream_t st, const Shape *sh = nullptr)
    {
        if (!sh) {
            GraphKey key{0, record, lo, hi, w.NS, w.seg, srcv, traces, nullptr};
            return replay(key, w.M, st, [&](hipStream_t s) { return run_forward_levels(w, lo, hi, srcv, traces, record, s); });
        }
        return run_forward_levels(w, lo, hi, srcv, traces, record, st, sh);
    }
    int run_forward_levels(const WS &w, int64_t lo, int64_t hi, const T *srcv, T *traces, int record,
                           hipStream_t st, const Shape *sh = nullptr)
    {
        const Layout L = plan_layout();
        for (int64_t n = lo + 1; n <= hi; n++) {
            int act = active[n];
            if (blocking == 2 && blk_ok && act > 0 && n + 1 <= hi && active[n + 1] == act) {
                ADS_TRY(fwd_pair(w, (int)n, act, srcv, traces, record, st, sh));
                n++;
                continue;
            }
            Lv<T> pi = phl(w, n - 1), si = psl(w, n - 1), po = phl(w, n), so = psl(w, n);
            ADS_TRY(fwd_step(L, (int)n, act, slot(w, n), cst(slot(w, n - 1)), cst(slot(w, n - 2)),
                             cst(pi), cst(si), po, so, w.C, srcv, traces, record, st, sh));
        }
        if (src_fused && srcv && record && traces && hi > lo && trace_elems > 0) {
            ac_record<T><<<dim3((unsigned)(hi - lo), (unsigned)nshots), 256, 0, st>>>(
                w.seg, w.NS, nshots * Npad, Npad, lo, d_nstep_sorted.as<int64_t>(), d_rcv_ptr.as<int64_t>(),
                d_rcv_cell[0].template as<int64_t>(), traces, d_trace_off.as<int64_t>());
            ADS_LAUNCH_CHECK();
        }
        return ADS_OK;
    }

    int frame_copy(bool unpack, Lv<T> f, T *rec, int act, hipStream_t st)
    {
        if (nfr <= 0 || act <= 0) return ADS_OK;
        const dim3 g(cdiv(nfr, 256), (unsigned)act);
        if (unpack) k_frame_copy<T, true><<<g, 256, 0, st>>>(f.p, Npad, rec, fbox, (int)Pitch);
        else k_frame_copy<T, false><<<g, 256, 0, st>>>(f.p, Npad, rec, fbox, (int)Pitch);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    // state at level a (u^a, u^{a-1}, phi^a, psi^a; active shots are a prefix) into slot q
    int save_ckpt(const WS &w, int64_t q, int64_t a, hipStream_t st)
    {
        const int act = rec_act(a);
        const size_t U = (size_t)act * Npad, R = (size_t)act * nfr;
        T *dst = w.ckpt + w.ck_off[q];
        ADS_HIP(hipMemcpyAsync(dst, slot(w, a).p, U * sizeof(T), hipMemcpyDeviceToDevice, st));
        ADS_HIP(hipMemcpyAsync(dst + U, slot(w, a - 1).p, U * sizeof(T), hipMemcpyDeviceToDevice, st));
        ADS_TRY(frame_copy(false, phl(w, a), dst + 2 * U, act, st));
        ADS_TRY(frame_copy(false, psl(w, a), dst + 2 * U + R, act, st));
        return ADS_OK;
    }

    // slot q (q < 0: the zero state at level 1) back into the slot ring and phi/psi buffers.
    // The buffers' quiet-box cells are zero: every forward clears them and no pass writes a
    // non-zero value there.
    int restore_ckpt(const WS &w, int64_t q, int64_t a, hipStream_t st)
    {
        const int act = rec_act(a);
        const size_t U = (size_t)act * Npad, R = (size_t)act * nfr;
        if (q < 0) {
            ADS_HIP(hipMemsetAsync(slot(w, 0).p, 0, U * sizeof(T), st));
            ADS_HIP(hipMemsetAsync(slot(w, 1).p, 0, U * sizeof(T), st));
            ADS_HIP(hipMemsetAsync(phl(w, 1).p, 0, U * sizeof(T), st));
            ADS_HIP(hipMemsetAsync(psl(w, 1).p, 0, U * sizeof(T), st));
            return ADS_OK;
        }
        const T *src = w.ckpt + w.ck_off[q];
        ADS_HIP(hipMemcpyAsync(slot(w, a).p, src, U * sizeof(T), hipMemcpyDeviceToDevice, st));
        ADS_HIP(hipMemcpyAsync(slot(w, a - 1).p, src + U, U * sizeof(T), hipMemcpyDeviceToDevice, st));
        ADS_TRY(frame_copy(true, phl(w, a), const_cast<T *>(src) + 2 * U, act, st));
        ADS_TRY(frame_copy(true, psl(w, a), const_cast<T *>(src) + 2 * U + R, act, st));
        return ADS_OK;
    }

    int check_ws(WS &w, void *ws, size_t wsb, int64_t K)
    {
        Carve cv(ws, wsb);
        layout(w, cv, K);
        if (!ws || !cv.ok())
            ADS_FAIL(ADS_ERR_WORKSPACE, "workspace too small: need %zu bytes, got %zu", cv.off, wsb);
        return ADS_OK;
    }

    int forward(const void *model, const void *srcv, void *traces, void *ws, size_t wsb,
                int64_t K, hipStream_t st)
    {
        WS w;
        ADS_TRY(check_ws(w, ws, wsb, K));
        const size_t F = (size_t)nshots * Npad * sizeof(T);
        ADS_TRY(prepare_C(w.C, model, (int)Pitch, st));
        ADS_HIP(hipMemsetAsync(slot(w, 0).p, 0, F, st));
        ADS_HIP(hipMemsetAsync(slot(w, 1).p, 0, F, st));
        for (int q = 0; q < 3; q++) {
            ADS_HIP(hipMemsetAsync(w.phi[q], 0, F, st));
            ADS_HIP(hipMemsetAsync(w.psi[q], 0, F, st));
        }
        if (traces && trace_elems) ADS_HIP(hipMemsetAsync(traces, 0, trace_elems * sizeof(T), st));
        // the reversal schedule's first descent: the states it keeps, in level order
        Reversal rv(w.M, w.S);
        const auto path = rv.first_path();
        size_t pi = 0;
        for (int64_t m = 0; m < w.M; m++) {
            const int64_t a = seg_level(w, m), b = seg_level(w, m + 1);
            if (pi < path.size() && path[pi].first == m) ADS_TRY(save_ckpt(w, path[pi++].second, a, st));
            ADS_TRY(run_forward(w, a, b, (const T *)srcv, (T *)traces, traces != nullptr, st));
        }
        return ADS_OK;
    }

    int zero_adjoint_state(const WS &w, hipStream_t st)
    {
        const size_t F = (size_t)nshots * Npad * sizeof(T);
        for (int q = 0; q < 2; q++) {
            ADS_HIP(hipMemsetAsync(w.P[q], 0, F, st));
            ADS_HIP(hipMemsetAsync(w.F[q], 0, F, st));
            ADS_HIP(hipMemsetAsync(w.G[q], 0, F, st));
        }
        ADS_HIP(hipMemsetAsync(w.gC, 0, (size_t)group_arrays() * Npad * sizeof(T), st));
        return ADS_OK;
    }

    int adj_at(const WS &w, int64_t n, const T *srcv, const T *resid, hipStream_t st,
               const Shape *sh = nullptr)
    {
        Lv<T> P{w.P[n & 1], Npad}, P1{w.P[(n + 1) & 1], Npad};
        Lv<T> F0{w.F[n & 1], Npad}, G0{w.G[n & 1], Npad};
        Lv<T> F1{w.F[(n + 1) & 1], Npad}, G1{w.G[(n + 1) & 1], Npad};
        return adj_level(plan_layout(), (int)n, active[n], P, cst(P1), F0, G0, cst(F1), cst(G1),
                         cst(slot(w, n - 1)), cst(slot(w, n)), Lv<const T>{nullptr, 0},
                         Lv<T>{w.gC, Npad}, w.C, srcv, resid, st, sh);
    }

    int backward(const void *model, const void *srcv, const T *resid, void *grad, void *ws,
                 size_t wsb, int64_t K, hipStream_t st)
    {
        WS w;
        ADS_TRY(check_ws(w, ws, wsb, K));
        ADS_TRY(zero_adjoint_state(w, st));
        const T *sv = (const T *)srcv;
        struct Ops {
            AcPlan *pl;
            const WS &w;
            const T *sv, *resid;
            hipStream_t st;
            int restore(int64_t m, int64_t q) { return pl->restore_ckpt(w, q, pl->seg_level(w, m), st); }
            int save(int64_t m, int64_t q) { return pl->save_ckpt(w, q, pl->seg_level(w, m), st); }
            int forward(int64_t m0, int64_t m1)  // segment by segment: one cached graph each
            {
                for (int64_t m = m0; m < m1; m++)
                    ADS_TRY(pl->run_forward(w, pl->seg_level(w, m), pl->seg_level(w, m + 1), sv, nullptr, 0, st));
                return ADS_OK;
            }
            int adjoint(int64_t m)
            {
                const int64_t a = pl->seg_level(w, m), b = pl->seg_level(w, m + 1);
                if (resid && pl->max_gi_cells > 0 && pl->inj_sg != pl->adj_shape.sg)
                    ADS_TRY(pl->build_inject_tables(pl->adj_shape.sg));  // host uploads: not in a capture
                GraphKey key{1, 0, a, b, w.NS, w.seg, sv, nullptr, resid};
                return pl->replay(key, w.M, st, [&](hipStream_t s) -> int {
                    for (int64_t n = b; n > a; n--) ADS_TRY(pl->adj_at(w, n, sv, resid, s));
                    return ADS_OK;
                });
            }
        } ops{this, w, sv, resid, st};
        Reversal rv(w.M, w.S);
        ADS_TRY(rv.run(ops));
        k_grad_out<T><<<dim3(cdiv(NY + 2, 256), NX + 2), 256, 0, st>>>(
            w.gC, Npad, (int)group_arrays(), (const T *)model, (T *)grad, (int)NX, (int)NY,
            (int)Pitch, p.square_velocity);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int misfit(const WS &w, const T *traces, const T *obs, hipStream_t st)
    {
        launch_misfit<T>(traces, obs, w.resid, trace_elems, w.part, w.loss, st);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    // ---- autotuner: time candidate (RPT, Sg, SB) shapes on the plan's own workload ----
    int autotune(const void *model, const void *srcv, void *ws, size_t wsb, int64_t K, int reps,
                 hipStream_t st)
    {
        drop_graphs();  // captured segments embed the kernel shapes
        WS w;
        ADS_TRY(check_ws(w, ws, wsb, K));
        if (NMAX < 3) {  // nothing worth timing
            tuned = 1;
            return ADS_OK;
        }
        const T *sv = (const T *)srcv;
        ADS_TRY(prepare_C(w.C, model, (int)Pitch, st));
        const size_t F = (size_t)nshots * Npad * sizeof(T);
        for (int64_t q = 0; q < std::min<int64_t>(w.NS, 4); q++) ADS_HIP(hipMemsetAsync(slot(w, q).p, 0, F, st));
        for (int q = 0; q < 3; q++) {
            ADS_HIP(hipMemsetAsync(w.phi[q], 0, F, st));
            ADS_HIP(hipMemsetAsync(w.psi[q], 0, F, st));
        }
        const int rpts[3] = {1, 2, 4};
        // the forward also tries the streaming two-level rectangle (8 rows per wavefront)
        const int frpts[4] = {1, 2, 4, 8};
        std::vector<int> sgs;
        for (int sg = 1; sg <= std::max<int64_t>(nshots, 1) && sg <= 32; sg *= 2) sgs.push_back(sg);
        if (reps < 1) reps = 3;
        hipEvent_t e0, e1;
        ADS_HIP(hipEventCreate(&e0));
        ADS_HIP(hipEventCreate(&e1));
        auto timeit = [&](auto &&body) -> float {
            body();  // warm
            (void)hipEventRecord(e0, st);
            for (int r = 0; r < reps; r++) body();
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            return ms;
        };
        int64_t n0 = std::min<int64_t>(2, NMAX);
        float best = 1e30f;
        Shape bf = fwd_shape;
        for (int rp : frpts)
            for (int sg : sgs)
                for (int sb = 1; sb <= std::min(2, sg); sb++) {
                    if (sb == 2 && rp >= 4) continue;
                    Shape sh{rp, sg, sb};
                    float ms = timeit([&] { (void)run_forward(w, n0 - 1, n0 + 1 <= NMAX ? n0 + 1 : n0, sv, nullptr, 0, st, &sh); });
                    if (ms < best) best = ms, bf = sh;
                }
        fwd_shape = bf;
        ADS_TRY(zero_adjoint_state(w, st));
        // the receiver injection's cost depends on the shot grouping: time it with the stencils
        if (trace_elems) ADS_HIP(hipMemsetAsync(w.resid, 0, (size_t)trace_elems * sizeof(T), st));
        best = 1e30f;
        Shape ba = adj_shape;
        const Shape keep = adj_shape;
        for (int rp : rpts)
            for (int sg : sgs)
                for (int sb = 1; sb <= std::min(2, sg); sb++) {
                    if (sb == 2 && rp == 4) continue;
                    Shape sh{rp, sg, sb};
                    adj_shape = sh;  // the gC grouping follows the shape being timed
                    float ms = timeit([&] { (void)adj_at(w, std::max<int64_t>(n0, 2), sv, w.resid, st, &sh); });
                    if (ms < best) best = ms, ba = sh;
                }
        adj_shape = nshots > 0 ? ba : keep;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        tuned = 1;
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    // ---- full-history drop-in (every shot has NSTEP levels, dense caller layout) ----
    int forward_history(const void *model, const void *srcv, void *u_hist, void *phi_hist,
                        void *psi_hist, void *scratch, hipStream_t st)
    {
        const Layout L = dense_layout();
        const int64_t Lv_ = NMAX + 1;
        const size_t H = (size_t)nshots * Lv_ * Ndense * sizeof(T);
        T *C = (T *)scratch;
        const size_t F = (size_t)nshots * Ndense;
        T *pp[2] = {(T *)scratch + Ndense, (T *)scratch + Ndense + F};
        T *qq[2] = {(T *)scratch + Ndense + 2 * F, (T *)scratch + Ndense + 3 * F};
        ADS_TRY(prepare_C(C, model, L.P, st));
        ADS_HIP(hipMemsetAsync(u_hist, 0, H, st));
        if (phi_hist) ADS_HIP(hipMemsetAsync(phi_hist, 0, H, st));
        if (psi_hist) ADS_HIP(hipMemsetAsync(psi_hist, 0, H, st));
        if (!phi_hist || !psi_hist)
            for (int q = 0; q < 2; q++) {
                ADS_HIP(hipMemsetAsync(pp[q], 0, F * sizeof(T), st));
                ADS_HIP(hipMemsetAsync(qq[q], 0, F * sizeof(T), st));
            }
        const int64_t hs = Lv_ * Ndense;
        auto lvl = [&](void *base, int64_t n) { return Lv<T>{(T *)base + n * Ndense, hs}; };
        for (int64_t n = 2; n <= NMAX; n++) {
            Lv<T> pi = phi_hist ? lvl(phi_hist, n - 1) : Lv<T>{pp[(n - 1) & 1], Ndense};
            Lv<T> po = phi_hist ? lvl(phi_hist, n) : Lv<T>{pp[n & 1], Ndense};
            Lv<T> si = psi_hist ? lvl(psi_hist, n - 1) : Lv<T>{qq[(n - 1) & 1], Ndense};
            Lv<T> so = psi_hist ? lvl(psi_hist, n) : Lv<T>{qq[n & 1], Ndense};
            ADS_TRY(fwd_step(L, (int)n, (int)nshots, lvl(u_hist, n), cst(lvl(u_hist, n - 1)),
                             cst(lvl(u_hist, n - 2)), cst(pi), cst(si), po, so, C,
                             (const T *)srcv, nullptr, 0, st));
        }
        return ADS_OK;
    }

    int backward_history(const void *model, const void *srcv, const void *u_hist,
                         const void *g_hist, void *grad, void *scratch, hipStream_t st)
    {
        const Layout L = dense_layout();
        const int64_t Lv_ = NMAX + 1;
        const size_t F = (size_t)nshots * Ndense;
        T *C = (T *)scratch;
        T *b = (T *)scratch + Ndense;
        T *P[2] = {b, b + F}, *Fa[2] = {b + 2 * F, b + 3 * F}, *Ga[2] = {b + 4 * F, b + 5 * F};
        T *gC = b + 6 * F;
        ADS_TRY(prepare_C(C, model, L.P, st));
        ADS_HIP(hipMemsetAsync(b, 0, 7 * F * sizeof(T), st));
        const int64_t hs = Lv_ * Ndense;
        auto lvl = [&](const void *base, int64_t n) {
            return Lv<const T>{(const T *)base + n * Ndense, hs};
        };
        for (int64_t n = NMAX; n >= 2; n--) {
            ADS_TRY(adj_level(L, (int)n, (int)nshots, Lv<T>{P[n & 1], Ndense},
                              Lv<const T>{P[(n + 1) & 1], Ndense}, Lv<T>{Fa[n & 1], Ndense},
                              Lv<T>{Ga[n & 1], Ndense}, Lv<const T>{Fa[(n + 1) & 1], Ndense},
                              Lv<const T>{Ga[(n + 1) & 1], Ndense}, lvl(u_hist, n - 1),
                              lvl(u_hist, n), lvl(g_hist, n), Lv<T>{gC, Ndense}, C,
                              (const T *)srcv, nullptr, st));
        }
        k_grad_out<T><<<dim3(cdiv(NY + 2, 256), NX + 2), 256, 0, st>>>(
            gC, Ndense, (int)group_arrays(), (const T *)model, (T *)grad, (int)NX, (int)NY, L.P,
            p.square_velocity);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }
    size_t history_scratch() const { return ((size_t)Ndense + 7 * (size_t)nshots * Ndense) * sizeof(T); }
};

template <typename T>
static int build_plan(const ads_acoustic_params *p, const ads_shot_geometry *g, AcPlan<T> *P)
{
    ADS_TRY(validate_params(p));
    if (!g || g->nshots < 1 || !g->src_ptr || !g->rcv_ptr)
        ADS_FAIL(ADS_ERR_ARG, "geometry: need nshots >= 1, src_ptr and rcv_ptr");
    P->p = *p;
    P->variant = p->PropagatorKernel == 0 ? 0 : 1;
    P->NX = p->NX;
    P->NY = p->NY;
    P->Pitch = ((p->NY + 2 + WAVE - 1) / WAVE) * WAVE;
    P->Npad = (p->NX + 2) * P->Pitch;
    // the stencil kernels address one field of one shot with 32-bit byte offsets
    if ((uint64_t)P->Npad * sizeof(T) >= (uint64_t(1) << 32))
        ADS_FAIL(ADS_ERR_ARG, "grid too large: one padded field of one shot must be below 4 GiB");
    P->Ndense = (p->NX + 2) * (p->NY + 2);
    P->nshots = g->nshots;
    P->k = make_coef<T>(p->DELTAT, p->DELTAX, p->DELTAY);
    // the benchmark shape (forward: the streaming two-level rectangle, 8 rows per wavefront;
    // adjoint: 2 rows per wavefront), with at most 16 shots per workgroup group
    P->fwd_shape = Shape{8, (int)std::min<int64_t>(16, g->nshots), 1};
    P->adj_shape = Shape{2, (int)std::min<int64_t>(16, g->nshots), 1};
    const int64_t ns = g->nshots;
    if (g->src_ptr[0] != 0 || g->rcv_ptr[0] != 0) ADS_FAIL(ADS_ERR_ARG, "src_ptr/rcv_ptr must start at 0");
    std::vector<int64_t> nstep(ns);
    for (int64_t s = 0; s < ns; s++) {
        nstep[s] = g->nstep ? g->nstep[s] : p->NSTEP;
        if (nstep[s] < 1 || nstep[s] > p->NSTEP)
            ADS_FAIL(ADS_ERR_SHAPE, "shot %lld: nstep %lld outside [1, NSTEP=%lld]", (long long)s,
                     (long long)nstep[s], (long long)p->NSTEP);
        if (g->src_ptr[s + 1] < g->src_ptr[s] || g->rcv_ptr[s + 1] < g->rcv_ptr[s])
            ADS_FAIL(ADS_ERR_ARG, "src_ptr/rcv_ptr must be non-decreasing");
    }
    // plan order: longest record first, so the shots active at level n are a prefix
    P->order.resize(ns);
    for (int64_t s = 0; s < ns; s++) P->order[s] = s;
    std::stable_sort(P->order.begin(), P->order.end(),
                     [&](int64_t a, int64_t b) { return nstep[a] > nstep[b]; });
    P->NMAX = nstep[P->order[0]];
    P->nstep_sorted.resize(ns);
    for (int64_t q = 0; q < ns; q++) P->nstep_sorted[q] = nstep[P->order[q]];
    P->active.assign(P->NMAX + 2, 0);
    for (int64_t n = 0; n <= P->NMAX + 1; n++) {
        int c = 0;
        for (int64_t q = 0; q < ns; q++) c += P->nstep_sorted[q] >= n;
        P->active[n] = c;
    }
    // caller-order offsets of srcv and traces
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
    // validated 1-based (i, j) in plan order
    auto check = [&](const int64_t *I, const int64_t *J, int64_t e, const char *what) -> int {
        int64_t i = I[e], j = J[e];
        if (i < 1 || i > p->NX + 2 || j < 1 || j > p->NY + 2)
            ADS_FAIL(ADS_ERR_INDEX, "%s %lld at (%lld, %lld) is outside the padded grid [1, %lld] x [1, %lld]",
                     what, (long long)e, (long long)i, (long long)j, (long long)(p->NX + 2),
                     (long long)(p->NY + 2));
        return ADS_OK;
    };
    std::vector<int64_t> sp(ns + 1, 0), rp(ns + 1, 0), si, sj, ri, rj, nsrc(ns), soff(ns), toff(ns);
    for (int64_t q = 0; q < ns; q++) {
        int64_t s = P->order[q];
        if (g->src_ptr[s + 1] > g->src_ptr[s] && (!g->srci || !g->srcj))
            ADS_FAIL(ADS_ERR_ARG, "srci/srcj are null");
        if (g->rcv_ptr[s + 1] > g->rcv_ptr[s] && (!g->rcvi || !g->rcvj))
            ADS_FAIL(ADS_ERR_ARG, "rcvi/rcvj are null");
        for (int64_t e = g->src_ptr[s]; e < g->src_ptr[s + 1]; e++) {
            ADS_TRY(check(g->srci, g->srcj, e, "source"));
            si.push_back(g->srci[e] - 1);
            sj.push_back(g->srcj[e] - 1);
        }
        for (int64_t e = g->rcv_ptr[s]; e < g->rcv_ptr[s + 1]; e++) {
            ADS_TRY(check(g->rcvi, g->rcvj, e, "receiver"));
            ri.push_back(g->rcvi[e] - 1);
            rj.push_back(g->rcvj[e] - 1);
        }
        sp[q + 1] = (int64_t)si.size();
        rp[q + 1] = (int64_t)ri.size();
        nsrc[q] = g->src_ptr[s + 1] - g->src_ptr[s];
        soff[q] = P->srcv_off_caller[s];
        toff[q] = P->trace_off_caller[s];
    }
    auto flat = [](const std::vector<int64_t> &I, const std::vector<int64_t> &J, int64_t pitch) {
        std::vector<int64_t> f(I.size());
        for (size_t e = 0; e < I.size(); e++) f[e] = I[e] * pitch + J[e];
        return f;
    };
    const int64_t pitches[2] = {P->Pitch, p->NY + 2};
    for (int L = 0; L < 2; L++) {
        std::vector<int64_t> sf = flat(si, sj, pitches[L]), rf = flat(ri, rj, pitches[L]);
        CellLists us = build_cell_lists(ns, sp.data(), sf);
        // cell lists are identical in structure for both pitches (the order of cells is the
        // same because (i, j) -> i*pitch + j is monotone for j < pitch)
        if (L == 0) {
            ADS_TRY(P->d_us_ptr.upload(us.shot_ptr));
            ADS_TRY(P->d_us_eptr.upload(us.ent_ptr));
            ADS_TRY(P->d_us_ent.upload(us.ent));
        }
        ADS_TRY(P->d_us_cell[L].upload(us.cell));
        ADS_TRY(P->d_rcv_cell[L].upload(rf));
    }
    // PML profiles and flags
    P->sx.resize(p->NX + 2);
    P->ty.resize(p->NY + 2);
    ADS_TRY(ads_acoustic_pml_profiles(p, P->sx.data(), P->ty.data()));
    P->fbox = quiet_box(P->sx, P->ty);
    P->nfr = P->fbox.cells();
    std::vector<T> sxT(P->sx.begin(), P->sx.end()), tyT(P->ty.begin(), P->ty.end());
    auto flags = [](const std::vector<T> &v, std::vector<uint8_t> &pm, std::vector<uint8_t> &nr) {
        size_t L = v.size();
        pm.assign(L, 0);
        nr.assign(L, 0);
        for (size_t q = 0; q < L; q++) pm[q] = v[q] != T(0);
        for (size_t q = 0; q < L; q++)
            nr[q] = pm[q] | (q > 0 ? pm[q - 1] : 0) | (q + 1 < L ? pm[q + 1] : 0);
    };
    flags(sxT, P->pmlx, P->nearx);
    flags(tyT, P->pmly, P->neary);
    {
        int64_t cnt = 0;
        for (int64_t i = 1; i <= p->NX; i++)
            for (int64_t j = 1; j <= p->NY; j++) cnt += (P->pmlx[i] | P->pmly[j]) != 0;
        P->f_pml = (double)cnt / ((double)p->NX * (double)p->NY);
    }
    ADS_TRY(P->d_sx.upload(sxT));
    ADS_TRY(P->d_ty.upload(tyT));
    for (int li = 0; li < 2; li++) {
        std::vector<int4> items;
        // plain rows/columns: interior and no PML formula needed
        std::vector<uint8_t> px(p->NX + 2, 0), py(p->NY + 2, 0);
        for (int64_t i = 1; i <= p->NX; i++) px[i] = !P->nearx[i];
        for (int64_t j = 1; j <= p->NY; j++) py[j] = !P->neary[j];
        split_grid(px, py, p->NX, p->NY, li == 0 ? AcPlan<T>::VECP : 1, P->rect[li], items);
        P->nitems[li] = (int)items.size();
        ADS_TRY(P->d_items[li].upload(items));
    }
    {
        const Rect &R = P->rect[0];
        P->rect1 = Rect{R.iL + 1, R.iR - 1, R.jL + 1, R.jR - 1};
        bool src_in = false;
        for (size_t e = 0; e < si.size(); e++)
            src_in |= si[e] >= R.iL && si[e] < R.iR && sj[e] >= R.jL && sj[e] < R.jR;
        P->blk_ok = !src_in && P->rect1.iR > P->rect1.iL && P->rect1.jR > P->rect1.jL;
        // sources inside the frame work: per source cell, its entries in (shot, source) order
        P->src_fused = !src_in;
        if (P->src_fused) {
            struct Ent {
                int64_t cell;
                int32_t shot;
                int64_t base, ns;
            };
            std::vector<Ent> ents;
            for (int64_t q = 0; q < ns; q++)
                for (int64_t e = sp[q]; e < sp[q + 1]; e++)
                    ents.push_back(Ent{si[e] * P->Pitch + sj[e], (int32_t)q, soff[q] + (e - sp[q]), nsrc[q]});
            std::stable_sort(ents.begin(), ents.end(), [](const Ent &x, const Ent &y) {
                return x.cell < y.cell || (x.cell == y.cell && x.shot < y.shot);
            });
            std::vector<int32_t> idx((size_t)P->Npad, -1), ptr, shot;
            std::vector<int64_t> base, nsv;
            for (size_t e = 0; e < ents.size(); e++) {
                if (e == 0 || ents[e].cell != ents[e - 1].cell) {
                    idx[(size_t)ents[e].cell] = (int32_t)ptr.size();
                    ptr.push_back((int32_t)e);
                }
                shot.push_back(ents[e].shot);
                base.push_back(ents[e].base);
                nsv.push_back(ents[e].ns);
            }
            ptr.push_back((int32_t)ents.size());
            ADS_TRY(P->d_sc_idx.upload(idx));
            ADS_TRY(P->d_sc_ptr.upload(ptr));
            if (!ents.empty()) {
                ADS_TRY(P->d_sc_shot.upload(shot));
                ADS_TRY(P->d_sc_base.upload(base));
                ADS_TRY(P->d_sc_ns.upload(nsv));
            }
            ADS_TRY(P->d_nstep_sorted.upload(P->nstep_sorted));
        }
        std::vector<int4> items;
        frame_items(P->blk_ok ? P->rect1 : Rect{1, 1, 1, 1}, p->NX, p->NY, items);
        P->nitems1 = P->blk_ok ? (int)items.size() : 0;
        if (P->blk_ok) ADS_TRY(P->d_items1.upload(items));
        if (P->blk_ok && P->src_fused && P->variant == 1) {
            for (int h = 0; h < FRAME2_H; h++) {
                frame_items(P->rect1, p->NX, p->NY, items, frame2_height(h), WAVE - 2);
                P->nitems2[h] = (int)items.size();
                ADS_TRY(P->d_items2[h].upload(items));
            }
        }
        if (P->src_fused && P->variant == 0) {
            frame_items(P->rect[0], p->NX, p->NY, items, FRPT, WAVE - 2);
            P->nitems_h = (int)items.size();
            ADS_TRY(P->d_items_h.upload(items));
            if (P->blk_ok) {
                frame_items(P->rect1, p->NX, p->NY, items, FRPT, WAVE - 2);
                P->nitems1h = (int)items.size();
                ADS_TRY(P->d_items1h.upload(items));
            }
        }
    }
    ADS_TRY(P->d_pmlx.upload(P->pmlx));
    ADS_TRY(P->d_pmly.upload(P->pmly));
    ADS_TRY(P->d_nearx.upload(P->nearx));
    ADS_TRY(P->d_neary.upload(P->neary));
    ADS_TRY(P->d_nsrc.upload(nsrc));
    ADS_TRY(P->d_srcv_off.upload(soff));
    ADS_TRY(P->d_rcv_ptr.upload(rp));
    ADS_TRY(P->d_trace_off.upload(toff));
    P->rcv_i0 = ri;
    P->rcv_j0 = rj;
    P->rcv_ptr_h = rp;
    ADS_TRY(P->build_inject_tables(P->adj_shape.sg));
    return ADS_OK;
}

}  // namespace ads

using namespace ads;

struct ads_acoustic_plan {
    std::unique_ptr<AcPlanBase> impl;
};

template <typename F> static int dispatch(ads_acoustic_plan *pl, F &&f)
{
    if (!pl || !pl->impl) ADS_FAIL(ADS_ERR_ARG, "null plan");
    if (pl->impl->dtype == ADS_F64) return f(*static_cast<AcPlan<double> *>(pl->impl.get()));
    return f(*static_cast<AcPlan<float> *>(pl->impl.get()));
}

extern "C" {

int ads_acoustic_plan_create(const ads_acoustic_params *p, const ads_shot_geometry *g, int dtype,
                             ads_acoustic_plan **out)
{
    clear_error();
    if (!out) ADS_FAIL(ADS_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (dtype != ADS_F32 && dtype != ADS_F64) ADS_FAIL(ADS_ERR_ARG, "dtype must be ADS_F32 or ADS_F64");
    auto pl = std::make_unique<ads_acoustic_plan>();
    int rc;
    if (dtype == ADS_F64) {
        auto P = std::make_unique<AcPlan<double>>();
        P->dtype = dtype;
        rc = build_plan<double>(p, g, P.get());
        pl->impl = std::move(P);
    } else {
        auto P = std::make_unique<AcPlan<float>>();
        P->dtype = dtype;
        rc = build_plan<float>(p, g, P.get());
        pl->impl = std::move(P);
    }
    if (rc != ADS_OK) return rc;
    *out = pl.release();
    return ADS_OK;
}

int ads_acoustic_plan_destroy(ads_acoustic_plan *plan)
{
    delete plan;
    return ADS_OK;
}

int ads_acoustic_plan_sizes(const ads_acoustic_plan *plan, int64_t *srcv_elems,
                            int64_t *trace_elems, int64_t *srcv_off, int64_t *trace_off)
{
    clear_error();
    if (!plan || !plan->impl) ADS_FAIL(ADS_ERR_ARG, "null plan");
    const AcPlanBase &b = *plan->impl;
    if (srcv_elems) *srcv_elems = b.srcv_elems;
    if (trace_elems) *trace_elems = b.trace_elems;
    for (int64_t s = 0; s < b.nshots; s++) {
        if (srcv_off) srcv_off[s] = b.srcv_off_caller[s];
        if (trace_off) trace_off[s] = b.trace_off_caller[s];
    }
    return ADS_OK;
}

size_t ads_acoustic_plan_workspace_size(const ads_acoustic_plan *plan, int64_t K)
{
    if (!plan || !plan->impl) return 0;
    if (plan->impl->dtype == ADS_F64) return static_cast<AcPlan<double> *>(plan->impl.get())->ws_size(K);
    return static_cast<AcPlan<float> *>(plan->impl.get())->ws_size(K);
}

int64_t ads_acoustic_plan_choose_interval(const ads_acoustic_plan *plan, size_t budget)
{
    if (!plan || !plan->impl) return 0;
    // One-level schedule: the fewest segments M whose interval fits, then the shortest
    // interval with that M: every segment but the last is recomputed, so K = ceil(levels / M)
    // minimises the recomputation ((M - 1) K levels) among the intervals that need M segments.
    const int64_t L = std::max<int64_t>(plan->impl->NMAX - 1, 1);
    for (int64_t M = 1; M <= L; M++) {
        const int64_t K = (L + M - 1) / M;
        const size_t need = plan->impl->dtype == ADS_F64
                                ? static_cast<AcPlan<double> *>(plan->impl.get())->ws_size(K, -1)
                                : static_cast<AcPlan<float> *>(plan->impl.get())->ws_size(K, -1);
        if (need <= budget) return K;
        if (K == 1) break;
    }
    return 0;
}

int ads_acoustic_plan_set_snapshots(ads_acoustic_plan *plan, int64_t snapshots)
{
    clear_error();
    if (!plan || !plan->impl) ADS_FAIL(ADS_ERR_ARG, "null plan");
    plan->impl->snapshots = snapshots < 0 ? -1 : snapshots;
    return ADS_OK;
}

int64_t ads_acoustic_plan_snapshots(const ads_acoustic_plan *plan)
{
    return plan && plan->impl ? plan->impl->snapshots : -1;
}

int ads_acoustic_plan_choose_schedule(ads_acoustic_plan *plan, size_t budget, int64_t *K, int64_t *S)
{
    clear_error();
    if (!plan || !plan->impl || !K || !S) ADS_FAIL(ADS_ERR_ARG, "null argument");
    const int rc = plan->impl->dtype == ADS_F64
                       ? static_cast<AcPlan<double> *>(plan->impl.get())->choose_schedule(budget, K, S)
                       : static_cast<AcPlan<float> *>(plan->impl.get())->choose_schedule(budget, K, S);
    if (rc != ADS_OK) ADS_FAIL(rc, "no checkpoint schedule fits %zu bytes", budget);
    plan->impl->snapshots = *S;
    return ADS_OK;
}

int ads_acoustic_plan_forward(ads_acoustic_plan *plan, const void *model, const void *srcv,
                              void *traces, void *ws, size_t wsb, int64_t K, ads_stream_t stream)
{
    clear_error();
    if (!model || !srcv) ADS_FAIL(ADS_ERR_ARG, "null model/srcv");
    return dispatch(plan, [&](auto &P) { return P.forward(model, srcv, traces, ws, wsb, K, S(stream)); });
}

int ads_acoustic_plan_backward(ads_acoustic_plan *plan, const void *model, const void *srcv,
                               const void *grad_traces, void *grad_model, void *ws, size_t wsb,
                               int64_t K, ads_stream_t stream)
{
    clear_error();
    if (!model || !srcv || !grad_traces || !grad_model) ADS_FAIL(ADS_ERR_ARG, "null argument");
    return dispatch(plan, [&](auto &P) {
        using R = typename std::remove_const<decltype(P.k.dt)>::type;
        return P.backward(model, srcv, (const R *)grad_traces, grad_model, ws, wsb, K, S(stream));
    });
}

int ads_acoustic_plan_gradient(ads_acoustic_plan *plan, const void *model, const void *srcv,
                               const void *observed, double *loss, void *grad_model, void *traces,
                               void *ws, size_t wsb, int64_t K, ads_stream_t stream)
{
    clear_error();
    if (!model || !srcv || !observed || !grad_model) ADS_FAIL(ADS_ERR_ARG, "null argument");
    return dispatch(plan, [&](auto &P) -> int {
        using R = typename std::remove_const<decltype(P.k.dt)>::type;
        typename std::remove_reference<decltype(P)>::type::WS w;
        ADS_TRY(P.check_ws(w, ws, wsb, K));
        // traces go to the caller's buffer if given, else into the residual area (reused)
        R *tr = traces ? (R *)traces : w.resid;
        ADS_TRY(P.forward(model, srcv, tr, ws, wsb, K, S(stream)));
        ADS_TRY(P.misfit(w, tr, (const R *)observed, S(stream)));
        if (loss) ADS_HIP(hipMemcpyAsync(loss, w.loss, sizeof(double), hipMemcpyDeviceToDevice, S(stream)));
        return P.backward(model, srcv, w.resid, grad_model, ws, wsb, K, S(stream));
    });
}

int ads_acoustic_plan_autotune(ads_acoustic_plan *plan, const void *model, const void *srcv,
                               void *ws, size_t wsb, int64_t K, int reps, int *shape_out,
                               ads_stream_t stream)
{
    clear_error();
    if (!model || !srcv) ADS_FAIL(ADS_ERR_ARG, "null argument");
    return dispatch(plan, [&](auto &P) -> int {
        ADS_TRY(P.autotune(model, srcv, ws, wsb, K, reps, S(stream)));
        if (shape_out) {
            const Shape *sh[2] = {&P.fwd_shape, &P.adj_shape};
            for (int q = 0; q < 2; q++) {
                shape_out[3 * q] = sh[q]->rpt;
                shape_out[3 * q + 1] = sh[q]->sg;
                shape_out[3 * q + 2] = sh[q]->sb;
            }
        }
        return ADS_OK;
    });
}

int ads_acoustic_plan_set_shape(ads_acoustic_plan *plan, const int *shape)
{
    clear_error();
    if (!shape) ADS_FAIL(ADS_ERR_ARG, "null shape");
    for (int q = 0; q < 2; q++) {
        const int rpt = shape[3 * q], sg = shape[3 * q + 1], sb = shape[3 * q + 2];
        if (rpt != 1 && rpt != 2 && rpt != 4 && !(q == 0 && (rpt == 8 || rpt == 16)))
            ADS_FAIL(ADS_ERR_ARG, "rows per wavefront must be 1, 2 or 4 (8 or 16 for the forward)");
        if (sg < 1) ADS_FAIL(ADS_ERR_ARG, "shots per workgroup must be >= 1");
        if (sb != 1 && sb != 2) ADS_FAIL(ADS_ERR_ARG, "shots per batch must be 1 or 2");
        if (sb == 2 && rpt >= 4) ADS_FAIL(ADS_ERR_ARG, "shots per batch 2 needs 1 or 2 rows per wavefront");
    }
    return dispatch(plan, [&](auto &P) -> int {
        P.fwd_shape = Shape{shape[0], shape[1], shape[2]};
        P.adj_shape = Shape{shape[3], shape[4], shape[5]};
        P.drop_graphs();
        return ADS_OK;
    });
}

int ads_acoustic_plan_set_graph(ads_acoustic_plan *plan, int enable)
{
    clear_error();
    return dispatch(plan, [&](auto &P) -> int {
        P.use_graph = enable ? 1 : 0;
        if (!enable) P.drop_graphs();
        return ADS_OK;
    });
}

int ads_acoustic_plan_set_frame_rows(ads_acoustic_plan *plan, int rows)
{
    clear_error();
    if (rows != 0 && rows != 4 && rows != 8 && rows != 16 && rows != 32 && rows != 64)
        ADS_FAIL(ADS_ERR_ARG, "frame rows must be 0, 4, 8, 16, 32 or 64, got %d", rows);
    return dispatch(plan, [&](auto &P) -> int {
        if (P.frame2_rows != rows) P.drop_graphs();
        P.frame2_rows = rows;
        return ADS_OK;
    });
}

int ads_acoustic_plan_set_blocking(ads_acoustic_plan *plan, int levels)
{
    clear_error();
    if (levels != 1 && levels != 2) ADS_FAIL(ADS_ERR_ARG, "levels must be 1 or 2, got %d", levels);
    return dispatch(plan, [&](auto &P) -> int {
        if (P.blocking != levels) P.drop_graphs();
        P.blocking = levels;
        return ADS_OK;
    });
}

int ads_acoustic_plan_blocking(const ads_acoustic_plan *plan)
{
    if (!plan || !plan->impl) return 0;
    return plan->impl->blocking == 2 && plan->impl->blk_ok ? 2 : 1;
}

int ads_acoustic_plan_forward_history(ads_acoustic_plan *plan, const void *model,
                                      const void *srcv, void *u_hist, void *phi_hist,
                                      void *psi_hist, void *scratch, ads_stream_t stream)
{
    clear_error();
    if (!model || !srcv || !u_hist || !scratch) ADS_FAIL(ADS_ERR_ARG, "null argument");
    return dispatch(plan, [&](auto &P) -> int {
        if (P.NMAX != P.p.NSTEP)
            ADS_FAIL(ADS_ERR_SHAPE, "history mode needs nstep == NSTEP for every shot");
        return P.forward_history(model, srcv, u_hist, phi_hist, psi_hist, scratch, S(stream));
    });
}

int ads_acoustic_plan_backward_history(ads_acoustic_plan *plan, const void *model,
                                       const void *srcv, const void *u_hist,
                                       const void *grad_u_hist, void *grad_model, void *scratch,
                                       ads_stream_t stream)
{
    clear_error();
    if (!model || !srcv || !u_hist || !grad_u_hist || !grad_model || !scratch)
        ADS_FAIL(ADS_ERR_ARG, "null argument");
    return dispatch(plan, [&](auto &P) -> int {
        if (P.NMAX != P.p.NSTEP)
            ADS_FAIL(ADS_ERR_SHAPE, "history mode needs nstep == NSTEP for every shot");
        return P.backward_history(model, srcv, u_hist, grad_u_hist, grad_model, scratch, S(stream));
    });
}

size_t ads_acoustic_plan_history_scratch_size(const ads_acoustic_plan *plan)
{
    if (!plan || !plan->impl) return 0;
    if (plan->impl->dtype == ADS_F64)
        return static_cast<AcPlan<double> *>(plan->impl.get())->history_scratch();
    return static_cast<AcPlan<float> *>(plan->impl.get())->history_scratch();
}

}  // extern "C"
