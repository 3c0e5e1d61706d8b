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
#include <vector>

#include "ads_common.h"

namespace ads {

constexpr int NF = 13;  // fields per level
enum { VX = 0, VY, SXX, SYY, SXY, M1, M2, M3, M4, M5, M6, M7, M8 };

// a level of the 13 arrays: field pointers and per-field shot strides
template <typename T> struct El {
    T *p[NF];
    int64_t st[NF];
    __device__ __host__ T *f(int q, int s) const { return p[q] + (int64_t)s * st[q]; }
};
template <typename T> struct ElC {
    const T *p[NF];
    int64_t st[NF];
    __device__ __host__ const T *f(int q, int s) const { return p[q] + (int64_t)s * st[q]; }
};
template <typename T> ElC<T> cst(const El<T> &a)
{
    ElC<T> c;
    for (int q = 0; q < NF; q++) c.p[q] = a.p[q], c.st[q] = a.st[q];
    return c;
}

template <typename T> struct ElCoef {
    T dt, ihx, ihy;  // 1/(24 hx), 1/(24 hy)
    const T *axf, *axh, *bxf, *bxh, *ayf, *ayh, *byf, *byh;
    int NX, NY, P;
};

// ------------------------------- forward -------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void el_stress_fwd(ElC<T> I, El<T> O, const T *__restrict__ lam,
                                                     const T *__restrict__ mu, ElCoef<T> c)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int s = blockIdx.z;
    const int NX = c.NX, NY = c.NY, P = c.P;
    if (j >= NY + 2 || i >= NX + 2) return;
    const int64_t ij = (int64_t)i * P + j;
    const T *__restrict__ vx = I.f(VX, s);
    const T *__restrict__ vy = I.f(VY, s);
    T sxx = I.f(SXX, s)[ij], syy = I.f(SYY, s)[ij], sxy = I.f(SXY, s)[ij];
    T m1 = 0, m2 = 0, m3 = 0, m4 = 0;
    if (i >= 1 && i <= NX - 1 && j >= 2 && j <= NY) {
        const T lb = T(0.5) * (lam[ij + P] + lam[ij]);
        const T mb = T(0.5) * (mu[ij + P] + mu[ij]);
        const T lm = lb + T(2) * mb;
        T dx = (T(27) * vx[ij + P] - T(27) * vx[ij] - vx[ij + 2 * P] + vx[ij - P]) * c.ihx;
        T dy = (T(27) * vy[ij] - T(27) * vy[ij - 1] - vy[ij + 1] + vy[ij - 2]) * c.ihy;
        m1 = c.bxh[i - 1] * I.f(M1, s)[ij] + c.axh[i - 1] * dx;
        m2 = c.byf[j - 1] * I.f(M2, s)[ij] + c.ayf[j - 1] * dy;
        dx += m1;
        dy += m2;
        sxx += (lm * dx + lb * dy) * c.dt;
        syy += (lm * dy + lb * dx) * c.dt;
    }
    if (i >= 2 && i <= NX && j >= 1 && j <= NY - 1) {
        const T mb = T(0.5) * (mu[ij] + mu[ij + 1]);
        T dx = (T(27) * vy[ij] - T(27) * vy[ij - P] - vy[ij + P] + vy[ij - 2 * P]) * c.ihx;
        T dy = (T(27) * vx[ij + 1] - T(27) * vx[ij] - vx[ij + 2] + vx[ij - 1]) * c.ihy;
        m3 = c.bxf[i - 1] * I.f(M3, s)[ij] + c.axf[i - 1] * dx;
        m4 = c.byh[j - 1] * I.f(M4, s)[ij] + c.ayh[j - 1] * dy;
        sxy += mb * (dx + m3 + dy + m4) * c.dt;
    }
    O.f(SXX, s)[ij] = sxx;
    O.f(SYY, s)[ij] = syy;
    O.f(SXY, s)[ij] = sxy;
    O.f(M1, s)[ij] = m1;
    O.f(M2, s)[ij] = m2;
    O.f(M3, s)[ij] = m3;
    O.f(M4, s)[ij] = m4;
}

template <typename T>
__global__ __launch_bounds__(256) void el_vel_fwd(ElC<T> I, El<T> O, const T *__restrict__ rho,
                                                  ElCoef<T> c)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int s = blockIdx.z;
    const int NX = c.NX, NY = c.NY, P = c.P;
    if (j >= NY + 2 || i >= NX + 2) return;
    const int64_t ij = (int64_t)i * P + j;
    const T *__restrict__ sxx = O.f(SXX, s);
    const T *__restrict__ syy = O.f(SYY, s);
    const T *__restrict__ sxy = O.f(SXY, s);
    T vx = I.f(VX, s)[ij], vy = I.f(VY, s)[ij];
    T m5 = 0, m6 = 0, m7 = 0, m8 = 0;
    if (i >= 2 && i <= NX && j >= 2 && j <= NY) {
        T dx = (T(27) * sxx[ij] - T(27) * sxx[ij - P] - sxx[ij + P] + sxx[ij - 2 * P]) * c.ihx;
        T dy = (T(27) * sxy[ij] - T(27) * sxy[ij - 1] - sxy[ij + 1] + sxy[ij - 2]) * c.ihy;
        m5 = c.bxf[i - 1] * I.f(M5, s)[ij] + c.axf[i - 1] * dx;
        m6 = c.byf[j - 1] * I.f(M6, s)[ij] + c.ayf[j - 1] * dy;
        vx += (dx + m5 + dy + m6) * c.dt / rho[ij];
    }
    if (i >= 1 && i <= NX - 1 && j >= 1 && j <= NY - 1) {
        const T rb = T(0.25) * (rho[ij] + rho[ij + P] + rho[ij + P + 1] + rho[ij + 1]);
        T dx = (T(27) * sxy[ij + P] - T(27) * sxy[ij] - sxy[ij + 2 * P] + sxy[ij - P]) * c.ihx;
        T dy = (T(27) * syy[ij + 1] - T(27) * syy[ij] - syy[ij + 2] + syy[ij - 1]) * c.ihy;
        m7 = c.bxh[i - 1] * I.f(M7, s)[ij] + c.axh[i - 1] * dx;
        m8 = c.byh[j - 1] * I.f(M8, s)[ij] + c.ayh[j - 1] * dy;
        vy += (dx + m7 + dy + m8) * c.dt / rb;
    }
    O.f(VX, s)[ij] = vx;
    O.f(VY, s)[ij] = vy;
    O.f(M5, s)[ij] = m5;
    O.f(M6, s)[ij] = m6;
    O.f(M7, s)[ij] = m7;
    O.f(M8, s)[ij] = m8;
}

// Sources at level n (unique (cell, type) keys, entries in source order), then receivers:
// trace[r * (nstep_s + 1) + n] = field_type_r[cell_r] (GetReceive.cpp:13-45 layout).
// mode 0: add; 1: add, then record the receivers; 2: remove the stress-type sources again
// (the adjoint needs the stresses of level n as fw3/fw4 read them, before injection).
template <typename T>
__global__ __launch_bounds__(256) void el_src_rcv(
    El<T> O, int n, const int64_t *__restrict__ us_ptr, const int64_t *__restrict__ us_key,
    const int64_t *__restrict__ us_eptr, const int64_t *__restrict__ us_ent,
    const T *__restrict__ srcv, const int64_t *__restrict__ srcv_off,
    const int64_t *__restrict__ nsrc, const int64_t *__restrict__ rcv_ptr,
    const int64_t *__restrict__ rcv_key, const int64_t *__restrict__ nstep,
    T *__restrict__ traces, const int64_t *__restrict__ trace_off, int mode)
{
    const int s = blockIdx.x;
    const int64_t ns = nsrc[s];
    const T *__restrict__ row = srcv + srcv_off[s] + (int64_t)(n - 1) * ns;
    for (int64_t c = us_ptr[s] + threadIdx.x; c < us_ptr[s + 1]; c += blockDim.x) {
        const int64_t key = us_key[c];
        const int type = (int)(key & 7);
        const int64_t cell = key >> 3;
        if (mode == 2 && type < SXX) continue;  // removal touches stresses only
        T *__restrict__ f = O.f(type, s);
        T acc = T(0);
        for (int64_t e = us_eptr[c]; e < us_eptr[c + 1]; e++) acc += row[us_ent[e]];
        f[cell] = mode == 2 ? f[cell] - acc : f[cell] + acc;
    }
    if (mode != 1) return;
    __syncthreads();
    const int64_t r0 = rcv_ptr[s], nr = rcv_ptr[s + 1] - r0, nt = nstep[s] + 1;
    T *__restrict__ tr = traces + trace_off[s];
    for (int64_t r = threadIdx.x; r < nr; r += blockDim.x) {
        const int64_t key = rcv_key[r0 + r];
        tr[r * nt + n] = O.f((int)(key & 7), s)[key >> 3];
    }
}

// ------------------------------- adjoint -------------------------------------------------
// cotangent injection of the trace residual at level n (unique (cell, type) per shot)
template <typename T>
__global__ __launch_bounds__(256) void el_adj_inject(
    El<T> A, int n, const int64_t *__restrict__ ur_ptr, const int64_t *__restrict__ ur_key,
    const int64_t *__restrict__ ur_eptr, const int64_t *__restrict__ ur_ent,
    const int64_t *__restrict__ nstep, const T *__restrict__ resid,
    const int64_t *__restrict__ trace_off)
{
    const int s = blockIdx.y;
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

// velocity phase, pass a: per-cell cotangents of the fw3/fw4 derivatives (T0..T3), the
// rho-bar cotangent of fw4 (T4), memory cotangents m5..m8, and fw3's d/drho.
template <typename T>
__global__ __launch_bounds__(256) void el_adj_vel_a(El<T> A, ElC<T> Sn, T *__restrict__ Tb,
                                                    int64_t tstride, int64_t shot_tstride,
                                                    T *__restrict__ grho, int64_t gstride,
                                                    const T *__restrict__ rho, ElCoef<T> c)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int s = blockIdx.z;
    const int NX = c.NX, NY = c.NY, P = c.P;
    if (j >= NY + 2 || i >= NX + 2) return;
    const int64_t ij = (int64_t)i * P + j;
    T *tb = Tb + (int64_t)s * shot_tstride;
    T t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    T *a5 = A.f(M5, s), *a6 = A.f(M6, s), *a7 = A.f(M7, s), *a8 = A.f(M8, s);
    T n5 = 0, n6 = 0, n7 = 0, n8 = 0;
    const T *__restrict__ sxx = Sn.f(SXX, s);
    const T *__restrict__ syy = Sn.f(SYY, s);
    const T *__restrict__ sxy = Sn.f(SXY, s);
    if (i >= 2 && i <= NX && j >= 2 && j <= NY) {
        const T dx = (T(27) * sxx[ij] - T(27) * sxx[ij - P] - sxx[ij + P] + sxx[ij - 2 * P]) * c.ihx;
        const T dy = (T(27) * sxy[ij] - T(27) * sxy[ij - 1] - sxy[ij + 1] + sxy[ij - 2]) * c.ihy;
        const T S3 = dx + Sn.f(M5, s)[ij] + dy + Sn.f(M6, s)[ij];
        const T g = A.f(VX, s)[ij], r = rho[ij];
        grho[(int64_t)s * gstride + ij] += -g * S3 * c.dt / (r * r);
        const T e = g * c.dt / r;
        const T g5 = a5[ij] + e, g6 = a6[ij] + e;
        t0 = e + c.axf[i - 1] * g5;
        t1 = e + c.ayf[j - 1] * g6;
        n5 = c.bxf[i - 1] * g5;
        n6 = c.byf[j - 1] * g6;
    }
    if (i >= 1 && i <= NX - 1 && j >= 1 && j <= NY - 1) {
        const T rb = T(0.25) * (rho[ij] + rho[ij + P] + rho[ij + P + 1] + rho[ij + 1]);
        const T dx = (T(27) * sxy[ij + P] - T(27) * sxy[ij] - sxy[ij + 2 * P] + sxy[ij - P]) * c.ihx;
        const T dy = (T(27) * syy[ij + 1] - T(27) * syy[ij] - syy[ij + 2] + syy[ij - 1]) * c.ihy;
        const T S4 = dx + Sn.f(M7, s)[ij] + dy + Sn.f(M8, s)[ij];
        const T g = A.f(VY, s)[ij];
        t4 = -g * S4 * c.dt / (rb * rb);
        const T e = g * c.dt / rb;
        const T g7 = a7[ij] + e, g8 = a8[ij] + e;
        t2 = e + c.axh[i - 1] * g7;
        t3 = e + c.ayh[j - 1] * g8;
        n7 = c.bxh[i - 1] * g7;
        n8 = c.byh[j - 1] * g8;
    }
    a5[ij] = n5;
    a6[ij] = n6;
    a7[ij] = n7;
    a8[ij] = n8;
    tb[ij] = t0;
    tb[tstride + ij] = t1;
    tb[2 * tstride + ij] = t2;
    tb[3 * tstride + ij] = t3;
    tb[4 * tstride + ij] = t4;
}

template <typename T> __device__ inline T tg(const T *t, int i, int j, int NX, int NY, int P)
{
    return (i >= 1 && i <= NX && j >= 1 && j <= NY) ? t[(int64_t)i * P + j] : T(0);
}

// velocity phase, pass b: transposed-stencil gathers into the stress cotangents and rho.
template <typename T>
__global__ __launch_bounds__(256) void el_adj_vel_b(El<T> A, const T *__restrict__ Tb, int64_t tstride,
                                                    int64_t shot_tstride, T *__restrict__ grho,
                                                    int64_t gstride, ElCoef<T> c)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int s = blockIdx.z;
    const int NX = c.NX, NY = c.NY, P = c.P;
    if (j >= NY + 2 || i >= NX + 2) return;
    const int64_t ij = (int64_t)i * P + j;
    const T *t0 = Tb + (int64_t)s * shot_tstride, *t1 = t0 + tstride, *t2 = t0 + 2 * tstride,
            *t3 = t0 + 3 * tstride, *t4 = t0 + 4 * tstride;
    // sxx <- (D-x)^T t0 ; sxy <- (D-y)^T t1 + (D+x)^T t2 ; syy <- (D+y)^T t3
    const T gxx = (T(27) * tg(t0, i, j, NX, NY, P) - T(27) * tg(t0, i + 1, j, NX, NY, P) -
                   tg(t0, i - 1, j, NX, NY, P) + tg(t0, i + 2, j, NX, NY, P)) * c.ihx;
    const T gxy = (T(27) * tg(t1, i, j, NX, NY, P) - T(27) * tg(t1, i, j + 1, NX, NY, P) -
                   tg(t1, i, j - 1, NX, NY, P) + tg(t1, i, j + 2, NX, NY, P)) * c.ihy +
                  (T(27) * tg(t2, i - 1, j, NX, NY, P) - T(27) * tg(t2, i, j, NX, NY, P) -
                   tg(t2, i - 2, j, NX, NY, P) + tg(t2, i + 1, j, NX, NY, P)) * c.ihx;
    const T gyy = (T(27) * tg(t3, i, j - 1, NX, NY, P) - T(27) * tg(t3, i, j, NX, NY, P) -
                   tg(t3, i, j - 2, NX, NY, P) + tg(t3, i, j + 1, NX, NY, P)) * c.ihy;
    A.f(SXX, s)[ij] += gxx;
    A.f(SXY, s)[ij] += gxy;
    A.f(SYY, s)[ij] += gyy;
    const T gr = T(0.25) * (tg(t4, i, j, NX, NY, P) + tg(t4, i - 1, j, NX, NY, P) +
                            tg(t4, i - 1, j - 1, NX, NY, P) + tg(t4, i, j - 1, NX, NY, P));
    grho[(int64_t)s * gstride + ij] += gr;
}

// stress phase, pass a
template <typename T>
__global__ __launch_bounds__(256) void el_adj_stress_a(El<T> A, ElC<T> Vm, ElC<T> Mn,
                                                       T *__restrict__ Tb, int64_t tstride,
                                                       int64_t shot_tstride,
                                                       const T *__restrict__ lam,
                                                       const T *__restrict__ mu, ElCoef<T> c)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int s = blockIdx.z;
    const int NX = c.NX, NY = c.NY, P = c.P;
    if (j >= NY + 2 || i >= NX + 2) return;
    const int64_t ij = (int64_t)i * P + j;
    T *tb = Tb + (int64_t)s * shot_tstride;
    const T *__restrict__ vx = Vm.f(VX, s);
    const T *__restrict__ vy = Vm.f(VY, s);
    T *a1 = A.f(M1, s), *a2 = A.f(M2, s), *a3 = A.f(M3, s), *a4 = A.f(M4, s);
    T t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0, n1 = 0, n2 = 0, n3 = 0, n4 = 0;
    if (i >= 1 && i <= NX - 1 && j >= 2 && j <= NY) {
        const T lb = T(0.5) * (lam[ij + P] + lam[ij]);
        const T mb = T(0.5) * (mu[ij + P] + mu[ij]);
        const T lm = lb + T(2) * mb;
        const T dx = (T(27) * vx[ij + P] - T(27) * vx[ij] - vx[ij + 2 * P] + vx[ij - P]) * c.ihx;
        const T dy = (T(27) * vy[ij] - T(27) * vy[ij - 1] - vy[ij + 1] + vy[ij - 2]) * c.ihy;
        const T X = dx + Mn.f(M1, s)[ij], Y = dy + Mn.f(M2, s)[ij];
        const T gx = A.f(SXX, s)[ij] * c.dt, gy = A.f(SYY, s)[ij] * c.dt;
        const T glm = X * gx + Y * gy;
        t4 = Y * gx + X * gy + glm;  // d/d lambda-bar
        t5 = T(2) * glm;             // d/d mu-bar (fw1)
        const T gX = lm * gx + lb * gy, gY = lb * gx + lm * gy;
        const T g1 = a1[ij] + gX, g2 = a2[ij] + gY;
        t0 = gX + c.axh[i - 1] * g1;
        t1 = gY + c.ayf[j - 1] * g2;
        n1 = c.bxh[i - 1] * g1;
        n2 = c.byf[j - 1] * g2;
    }
    if (i >= 2 && i <= NX && j >= 1 && j <= NY - 1) {
        const T mb = T(0.5) * (mu[ij] + mu[ij + 1]);
        const T dx = (T(27) * vy[ij] - T(27) * vy[ij - P] - vy[ij + P] + vy[ij - 2 * P]) * c.ihx;
        const T dy = (T(27) * vx[ij + 1] - T(27) * vx[ij] - vx[ij + 2] + vx[ij - 1]) * c.ihy;
        const T g = A.f(SXY, s)[ij];
        t6 = g * (dx + Mn.f(M3, s)[ij] + dy + Mn.f(M4, s)[ij]) * c.dt;  // d/d mu-bar (fw2)
        const T gs = g * mb * c.dt;
        const T g3 = a3[ij] + gs, g4 = a4[ij] + gs;
        t2 = gs + c.axf[i - 1] * g3;
        t3 = gs + c.ayh[j - 1] * g4;
        n3 = c.bxf[i - 1] * g3;
        n4 = c.byh[j - 1] * g4;
    }
    a1[ij] = n1;
    a2[ij] = n2;
    a3[ij] = n3;
    a4[ij] = n4;
    tb[ij] = t0;
    tb[tstride + ij] = t1;
    tb[2 * tstride + ij] = t2;
    tb[3 * tstride + ij] = t3;
    tb[4 * tstride + ij] = t4;
    tb[5 * tstride + ij] = t5;
    tb[6 * tstride + ij] = t6;
}

// stress phase, pass b: gathers into the velocity cotangents, lambda and mu.
template <typename T>
__global__ __launch_bounds__(256) void el_adj_stress_b(El<T> A, const T *__restrict__ Tb,
                                                       int64_t tstride, int64_t shot_tstride,
                                                       T *__restrict__ glam, T *__restrict__ gmu,
                                                       int64_t gstride, ElCoef<T> c)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int s = blockIdx.z;
    const int NX = c.NX, NY = c.NY, P = c.P;
    if (j >= NY + 2 || i >= NX + 2) return;
    const int64_t ij = (int64_t)i * P + j;
    const T *t0 = Tb + (int64_t)s * shot_tstride, *t1 = t0 + tstride, *t2 = t0 + 2 * tstride,
            *t3 = t0 + 3 * tstride, *t4 = t0 + 4 * tstride, *t5 = t0 + 5 * tstride,
            *t6 = t0 + 6 * tstride;
    // vx <- (D+x)^T t0 + (D+y)^T t3 ; vy <- (D-y)^T t1 + (D-x)^T t2
    const T gvx = (T(27) * tg(t0, i - 1, j, NX, NY, P) - T(27) * tg(t0, i, j, NX, NY, P) -
                   tg(t0, i - 2, j, NX, NY, P) + tg(t0, i + 1, j, NX, NY, P)) * c.ihx +
                  (T(27) * tg(t3, i, j - 1, NX, NY, P) - T(27) * tg(t3, i, j, NX, NY, P) -
                   tg(t3, i, j - 2, NX, NY, P) + tg(t3, i, j + 1, NX, NY, P)) * c.ihy;
    const T gvy = (T(27) * tg(t1, i, j, NX, NY, P) - T(27) * tg(t1, i, j + 1, NX, NY, P) -
                   tg(t1, i, j - 1, NX, NY, P) + tg(t1, i, j + 2, NX, NY, P)) * c.ihy +
                  (T(27) * tg(t2, i, j, NX, NY, P) - T(27) * tg(t2, i + 1, j, NX, NY, P) -
                   tg(t2, i - 1, j, NX, NY, P) + tg(t2, i + 2, j, NX, NY, P)) * c.ihx;
    A.f(VX, s)[ij] += gvx;
    A.f(VY, s)[ij] += gvy;
    const int64_t g = (int64_t)s * gstride + ij;
    glam[g] += T(0.5) * (tg(t4, i, j, NX, NY, P) + tg(t4, i - 1, j, NX, NY, P));
    gmu[g] += T(0.5) * (tg(t5, i, j, NX, NY, P) + tg(t5, i - 1, j, NX, NY, P)) +
              T(0.5) * (tg(t6, i, j, NX, NY, P) + tg(t6, i, j - 1, NX, NY, P));
}

template <typename T>
__global__ void el_model_in(const T *__restrict__ m, T *__restrict__ o, int NX, int NY, int P)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= NY + 2) return;
    o[(int64_t)i * P + j] = m[(int64_t)i * (NY + 2) + j];
}

template <typename T>
__global__ void el_grad_out(const T *__restrict__ g, int64_t stride, int nshots, T *__restrict__ out,
                            int NX, int NY, int P)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= NY + 2) return;
    const int64_t q = (int64_t)i * P + j;
    T acc = T(0);
    for (int s = 0; s < nshots; s++) acc += g[(int64_t)s * stride + q];
    out[(int64_t)i * (NY + 2) + j] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void el_misfit(const T *__restrict__ tr, const T *__restrict__ obs,
                                                 T *__restrict__ resid, int64_t n,
                                                 double *__restrict__ loss)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        T d = tr[q] - obs[q];
        resid[q] = T(2) * d;
        acc += (double)d * (double)d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
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
        c.axh = b + NX;
        c.bxf = b + 2 * NX;
        c.bxh = b + 3 * NX;
        c.ayf = b + 4 * NX;
        c.ayh = b + 4 * NX + NY;
        c.byf = b + 4 * NX + 2 * NY;
        c.byh = b + 4 * NX + 3 * NY;
        c.NX = (int)NX;
        c.NY = (int)NY;
        c.P = P;
        return c;
    }
    dim3 grid(int act) const { return dim3(cdiv(NY + 2, 64), cdiv(NX + 2, 4), act); }

    struct WS {
        int64_t K, NS, M;
        T *seg, *ckpt, *adj, *tmp, *grad, *rho, *lam, *mu, *resid;
        double *loss;
    };
    int64_t seg_count(int64_t K) const { return NMAX >= 1 ? (NMAX + K - 1) / K : 0; }
    void layout(WS &w, Carve &cv, int64_t K) const
    {
        if (K < 1) K = 1;
        if (K > std::max<int64_t>(NMAX, 1)) K = std::max<int64_t>(NMAX, 1);
        w.K = K;
        w.M = seg_count(K);
        w.NS = std::min<int64_t>(K + 1, NMAX + 1);
        const size_t F = (size_t)nshots * Npad * sizeof(T);
        w.seg = (T *)cv.take(F * NF * w.NS);
        w.ckpt = (T *)cv.take(F * NF * (size_t)std::max<int64_t>(w.M - 1, 0));
        w.adj = (T *)cv.take(F * NF);
        w.tmp = (T *)cv.take(F * 7);
        w.grad = (T *)cv.take(F * 3);
        w.rho = (T *)cv.take((size_t)Npad * sizeof(T));
        w.lam = (T *)cv.take((size_t)Npad * sizeof(T));
        w.mu = (T *)cv.take((size_t)Npad * sizeof(T));
        w.resid = (T *)cv.take((size_t)std::max<int64_t>(trace_elems, 1) * sizeof(T));
        w.loss = (double *)cv.take(64);
    }
    size_t ws_size(int64_t K) const
    {
        WS w;
        Carve cv(nullptr, 0);
        layout(w, cv, K);
        return cv.off + 256;
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
    El<T> slot(const WS &w, int64_t n) const
    {
        return block(w.seg + (size_t)(n % w.NS) * NF * nshots * Npad);
    }

    int step(const El<T> &I, const El<T> &O, int n, int act, const T *rho, const T *lam,
             const T *mu, const T *srcv, T *traces, int record, int P, int li, hipStream_t st)
    {
        if (act <= 0) return ADS_OK;
        const ElCoef<T> c = coef(P);
        const double cells = (double)act * (double)NX * (double)NY;
        // algorithmic bytes: stress kernel reads v (2) + sigma (3) + lam, mu (2), writes sigma (3),
        // memory m1..m4 read+write in their strips; velocity kernel reads sigma (3) + v (2) + rho,
        // writes v (2), m5..m8 in their strips
        int pr = prof_start(ADS_PROF_EL_FWD_STRESS, st);
        el_stress_fwd<T><<<grid(act), 256, 0, st>>>(cst(I), O, lam, mu, c);
        prof_stop(pr, st, cells * sizeof(T) * (10.0 + 4.0 * f_x + 4.0 * f_y));
        pr = prof_start(ADS_PROF_EL_FWD_VEL, st);
        el_vel_fwd<T><<<grid(act), 256, 0, st>>>(cst(I), O, rho, c);
        prof_stop(pr, st, cells * sizeof(T) * (8.0 + 4.0 * f_x + 4.0 * f_y));
        el_src_rcv<T><<<act, 256, 0, st>>>(O, n, d_us_ptr.as<int64_t>(), d_us_key[li].template as<int64_t>(),
                                           d_us_eptr.as<int64_t>(), d_us_ent.as<int64_t>(), srcv,
                                           d_srcv_off.as<int64_t>(), d_nsrc.as<int64_t>(),
                                           d_rcv_ptr.as<int64_t>(), d_rcv_key[li].template as<int64_t>(),
                                           d_nstep.as<int64_t>(), traces,
                                           d_trace_off.as<int64_t>(), record ? 1 : 0);
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
                         traces, record, (int)Pitch, 0, st));
        return ADS_OK;
    }

    int copy_state(const El<T> &dst, const El<T> &src, int act, hipStream_t st)
    {
        for (int q = 0; q < NF; q++)
            ADS_HIP(hipMemcpyAsync(dst.p[q], src.p[q], (size_t)act * Npad * sizeof(T),
                                   hipMemcpyDeviceToDevice, st));
        return ADS_OK;
    }
    El<T> ckpt(const WS &w, int64_t m) const
    {
        return block(w.ckpt + (size_t)(m - 1) * NF * nshots * Npad);
    }

    int forward(const void *rho, const void *lam, const void *mu, const void *srcv, void *traces,
                void *ws, size_t wsb, int64_t K, hipStream_t st)
    {
        WS w;
        ADS_TRY(check_ws(w, ws, wsb, K));
        ADS_TRY(prepare_models(w.rho, w.lam, w.mu, rho, lam, mu, st));
        ADS_HIP(hipMemsetAsync(slot(w, 0).p[0], 0, (size_t)NF * nshots * Npad * sizeof(T), st));
        if (traces && trace_elems) ADS_HIP(hipMemsetAsync(traces, 0, trace_elems * sizeof(T), st));
        for (int64_t m = 0; m < w.M; m++) {
            int64_t a = m * w.K, b = std::min<int64_t>(a + w.K, NMAX);
            if (m > 0) ADS_TRY(copy_state(ckpt(w, m), slot(w, a), active[std::min<int64_t>(a + 1, NMAX)], st));
            ADS_TRY(run_forward(w, a, b, (const T *)srcv, (T *)traces, traces != nullptr, st));
        }
        return ADS_OK;
    }

    // one reverse level: cotangents of level n -> level n-1.  Sn: state of level n, Sm: n-1.
    int adj_level(const El<T> &A, const El<T> &Sn, const El<T> &Sm, int n, int act,
                  const T *resid, const ElC<T> *dense, T *tmp, T *grad, const T *rho,
                  const T *lam, const T *mu, const T *srcv, int P, int li, hipStream_t st)
    {
        if (act <= 0) return ADS_OK;
        const ElCoef<T> c = coef(P);
        if (resid && max_ur > 0) {
            dim3 g(cdiv(max_ur, 256), act);
            el_adj_inject<T><<<g, 256, 0, st>>>(A, n, d_ur_ptr.as<int64_t>(), d_ur_key[li].template as<int64_t>(),
                                                d_ur_eptr.as<int64_t>(), d_ur_ent.as<int64_t>(),
                                                d_nstep.as<int64_t>(), resid, d_trace_off.as<int64_t>());
        }
        if (dense) el_adj_dense<T><<<grid(act), 256, 0, st>>>(A, *dense, (int)NX, (int)NY, P);
        el_src_rcv<T><<<act, 256, 0, st>>>(Sn, n, d_us_ptr.as<int64_t>(), d_us_key[li].template as<int64_t>(),
                                           d_us_eptr.as<int64_t>(), d_us_ent.as<int64_t>(), srcv,
                                           d_srcv_off.as<int64_t>(), d_nsrc.as<int64_t>(),
                                           d_rcv_ptr.as<int64_t>(), d_rcv_key[li].template as<int64_t>(),
                                           d_nstep.as<int64_t>(), nullptr,
                                           d_trace_off.as<int64_t>(), 2);
        const int64_t F = (int64_t)nshots * Npad;  // temporaries / gradients: [k][shot][Npad]
        const double cells = (double)act * (double)NX * (double)NY;
        T *grho = grad, *glam = grad + F, *gmu = grad + 2 * F;
        int pr = prof_start(ADS_PROF_EL_ADJ_VEL, st);
        el_adj_vel_a<T><<<grid(act), 256, 0, st>>>(A, cst(Sn), tmp, F, Npad, grho, Npad, rho, c);
        el_adj_vel_b<T><<<grid(act), 256, 0, st>>>(A, tmp, F, Npad, grho, Npad, c);
        // velocity VJP: reads v-bar (2), sigma (3), rho, stress cotangents r+w (6), rho grad r+w (2),
        // memory values and cotangents in the strips
        prof_stop(pr, st, cells * sizeof(T) * (14.0 + 6.0 * (f_x + f_y)));
        pr = prof_start(ADS_PROF_EL_ADJ_STRESS, st);
        el_adj_stress_a<T><<<grid(act), 256, 0, st>>>(A, cst(Sm), cst(Sn), tmp, F, Npad, lam, mu, c);
        el_adj_stress_b<T><<<grid(act), 256, 0, st>>>(A, tmp, F, Npad, glam, gmu, Npad, c);
        prof_stop(pr, st, cells * sizeof(T) * (14.0 + 6.0 * (f_x + f_y)));
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
        ADS_HIP(hipMemsetAsync(w.grad, 0, F * 3, st));
        ADS_HIP(hipMemsetAsync(w.tmp, 0, F * 7, st));
        const T *sv = (const T *)srcv;
        const El<T> A = block(w.adj);
        const int64_t L = NMAX + 1;
        for (int64_t m = w.M - 1; m >= 0; m--) {
            int64_t a = m * w.K, b = std::min<int64_t>(a + w.K, NMAX);
            if (m < w.M - 1) {
                int act = active[std::min<int64_t>(a + 1, NMAX)];
                if (m == 0) {
                    El<T> z = slot(w, 0);
                    for (int q = 0; q < NF; q++)
                        ADS_HIP(hipMemsetAsync(z.p[q], 0, (size_t)act * Npad * sizeof(T), st));
                } else {
                    ADS_TRY(copy_state(slot(w, a), ckpt(w, m), act, st));
                }
                ADS_TRY(run_forward(w, a, b, sv, nullptr, 0, st));
            }
            for (int64_t n = b; n > a; n--) {
                ElC<T> D, *Dp = nullptr;
                if (dense_hist) {  // [shot][field][level][Ndense]
                    for (int q = 0; q < 5; q++) {
                        D.p[q] = (const T *)dense_hist + ((size_t)q * L + n) * Ndense;
                        D.st[q] = 5 * L * Ndense;
                    }
                    for (int q = 5; q < NF; q++) D.p[q] = nullptr, D.st[q] = 0;
                    Dp = &D;
                }
                ADS_TRY(adj_level(A, slot(w, n), slot(w, n - 1), (int)n, active[n], resid, Dp,
                                  w.tmp, w.grad, w.rho, w.lam, w.mu, sv, (int)Pitch, 0, st));
            }
        }
        dim3 g(cdiv(NY + 2, 256), NX + 2);
        const int64_t gs = (int64_t)nshots * Npad;
        el_grad_out<T><<<g, 256, 0, st>>>(w.grad, Npad, (int)nshots, (T *)grho, (int)NX, (int)NY, (int)Pitch);
        el_grad_out<T><<<g, 256, 0, st>>>(w.grad + gs, Npad, (int)nshots, (T *)glam, (int)NX, (int)NY, (int)Pitch);
        el_grad_out<T><<<g, 256, 0, st>>>(w.grad + 2 * gs, Npad, (int)nshots, (T *)gmu, (int)NX, (int)NY, (int)Pitch);
        ADS_LAUNCH_CHECK();
        return ADS_OK;
    }

    int misfit(const WS &w, const T *traces, const T *obs, hipStream_t st)
    {
        ADS_HIP(hipMemsetAsync(w.loss, 0, sizeof(double), st));
        unsigned nb = std::min<unsigned>(cdiv(trace_elems, 256), 2048u);
        if (nb == 0) nb = 1;
        el_misfit<T><<<nb, 256, 0, st>>>(traces, obs, w.resid, trace_elems, w.loss);
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
    if ((p->NX + 2) * (((p->NY + 2 + 63) / 64) * 64) >= (int64_t)1 << 31 || p->NX + 2 > 65535)
        ADS_FAIL(ADS_ERR_SHAPE, "grid too large");
    if (!(p->DELTAT > 0) || !(p->DELTAX > 0) || !(p->DELTAY > 0))
        ADS_FAIL(ADS_ERR_ARG, "DELTAT, DELTAX, DELTAY must be > 0");
    if (p->subdomain)
        ADS_FAIL(ADS_ERR_ARG, "the MPIElastic subdomain variant is not implemented by the elastic plan");
    return ADS_OK;
}

template <typename T>
static int build_el_plan(const ads_elastic_params *p, const ads_shot_geometry *g, ElPlan<T> *P)
{
    ADS_TRY(validate_params(p));
    if (!g || g->nshots < 1 || !g->src_ptr || !g->rcv_ptr)
        ADS_FAIL(ADS_ERR_ARG, "geometry: need nshots >= 1, src_ptr and rcv_ptr");
    P->p = *p;
    P->NX = p->NX;
    P->NY = p->NY;
    P->Pitch = ((p->NY + 2 + 63) / 64) * 64;
    P->Npad = (p->NX + 2) * P->Pitch;
    P->Ndense = (p->NX + 2) * (p->NY + 2);
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
    auto check = [&](const int64_t *I, const int64_t *J, const int64_t *Ty, int64_t e, const char *what) -> int {
        if (!Ty) ADS_FAIL(ADS_ERR_ARG, "%s types are null", what);
        int64_t i = I[e], j = J[e], t = Ty[e];
        if (i < 1 || i > p->NX + 2 || j < 1 || j > p->NY + 2)
            ADS_FAIL(ADS_ERR_INDEX, "%s %lld at (%lld, %lld) is outside the padded grid [1, %lld] x [1, %lld]",
                     what, (long long)e, (long long)i, (long long)j, (long long)(p->NX + 2), (long long)(p->NY + 2));
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
            ri.push_back(g->rcvi[e] - 1);
            rj.push_back(g->rcvj[e] - 1);
            rt.push_back(g->rcvtype[e]);
        }
        sp[q + 1] = (int64_t)si.size();
        rp[q + 1] = (int64_t)ri.size();
        nsrc[q] = g->src_ptr[s + 1] - g->src_ptr[s];
        soff[q] = P->srcv_off_caller[s];
        toff[q] = P->trace_off_caller[s];
        nsv[q] = nstep[s];
    }
    const int64_t pitches[2] = {P->Pitch, p->NY + 2};
    for (int L = 0; L < 2; L++) {
        std::vector<int64_t> skey(si.size()), rkey(ri.size());
        for (size_t e = 0; e < si.size(); e++) skey[e] = (si[e] * pitches[L] + sj[e]) * 8 + st_[e];
        for (size_t e = 0; e < ri.size(); e++) rkey[e] = (ri[e] * pitches[L] + rj[e]) * 8 + rt[e];
        CellLists us = build_cell_lists(ns, sp.data(), skey), ur = build_cell_lists(ns, rp.data(), rkey);
        if (L == 0) {
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
    ADS_TRY(P->d_prof.upload(prof));
    int64_t cx = 0, cy = 0;
    for (int64_t k = 0; k < p->NX; k++) cx += (ax[k] != 0.0 || ax[p->NX + k] != 0.0);
    for (int64_t k = 0; k < p->NY; k++) cy += (ay[k] != 0.0 || ay[p->NY + k] != 0.0);
    P->f_x = (double)cx / (double)p->NX;
    P->f_y = (double)cy / (double)p->NY;
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
    int64_t hi = std::max<int64_t>(plan->impl->NMAX, 1), best = 0;
    for (int64_t K = hi; K >= 1; K = (K > 64 ? K - std::max<int64_t>(K / 64, 1) : K - 1))
        if (ads_elastic_plan_workspace_size(plan, K) <= budget) {
            best = K;
            break;
        }
    return best;
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
