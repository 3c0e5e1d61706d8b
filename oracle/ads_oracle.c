/*
 * ads_oracle.c -- FP64 CPU restatement of ADSeismic.jl's wave-propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP library and the
 * CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  Nothing in the product path links, calls or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * ADSeismic.jl v0.5.0 tree).  Layout everywhere: padded grid (NX+2) x (NY+2), row-major,
 * flat index IJ = i*(NY+2)+j with i = x (slow) and j = y/depth (fast), 0-based.
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   or_acoustic_step (kernel 1)      : golden vectors from the reference's own
 *                                      AcousticOneStepCpu.h compiled by oracle/Makefile.
 *   or_acoustic_step_vjp (kernel 1)  : same golden vectors (the header's backward).
 *   or_acoustic_step (kernel 0)      : equals kernel 1 when sigma=tau=0 (phi=psi=0);
 *                                      closed-form homogeneous checks; FD + dot-product.
 *   or_acoustic_step_vjp (kernel 0)  : dot-product test against finite differences.
 *   or_acoustic_pml_axis             : closed form spot values (edge value = xi, 0 inside).
 *   or_acoustic_forward/_adjoint     : linearity in srcv, causality (first arrival time),
 *                                      central finite differences of the misfit.
 *   or_elastic_*                     : homogeneous P-wave arrival time, mirror symmetry of an
 *                                      explosive source, FD + dot-product for rho/lambda/mu.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

#ifdef _OPENMP
#include <omp.h>
#endif

/* ----------------------------------------------------------------------------------------
 * a2: acoustic PML profile.  pml_helper, src/Core.jl:738-761; damping, Core.jl:625-629.
 * sigma depends on x only and tau on y only, so one 1-D axis is evaluated at a time.
 * coord(k) = k*h + origin for k = 0..len-1; the right edge sits at h*(NXg+1) - L where NXg
 * is the global interior size (MPIAcoustic.jl:128-151 uses the same helper with an origin).
 * -------------------------------------------------------------------------------------- */
void or_acoustic_pml_axis(i64 len, double h, double origin, i64 NXg, i64 npml, double Rcoef,
                          double vp_ref, int use_min, int use_max, double *out)
{
    const double L = (double)npml * h;
    const double xi = vp_ref / L * log(1.0 / Rcoef);
    const double right = h * (double)(NXg + 1) - L;
    for (i64 k = 0; k < len; k++) {
        double x = (double)k * h + origin;
        double v = 0.0;
        if (x < L && use_min) {
            double d = fabs(L - x);
            v = xi * (d / L - sin(2.0 * M_PI * d / L) / (2.0 * M_PI));
        } else if (x > right && use_max) {
            double d = fabs(x - right);
            v = xi * (d / L - sin(2.0 * M_PI * d / L) / (2.0 * M_PI));
        }
        out[k] = v;
    }
}

/* ----------------------------------------------------------------------------------------
 * a3/a4: one acoustic step.
 *   kernel 1 : deps/CustomOps/AcousticOneStepCpu/AcousticOneStepCpu.h:1-48 (phi/psi from w)
 *   kernel 0 : src/Core.jl:528-549 (phi/psi from the new u, before source injection)
 * sigma/tau are full (NX+2)(NY+2) grids as the reference op takes them.  Ring cells are 0.
 * -------------------------------------------------------------------------------------- */
void or_acoustic_step(int kernel, const double *w, const double *wold, const double *phi,
                      const double *psi, const double *sigma, const double *tau, const double *c,
                      double dt, double hx, double hy, i64 NX, i64 NY, double *u, double *phiout,
                      double *psiout)
{
    const i64 S = NY + 2;
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < NX + 2; i++) {
        for (i64 j = 0; j < NY + 2; j++) {
            i64 IJ = i * S + j;
            if (i == 0 || i == NX + 1 || j == 0 || j == NY + 1) {
                u[IJ] = 0.0;
                continue;
            }
            i64 IpJ = IJ + S, InJ = IJ - S, IJp = IJ + 1, IJn = IJ - 1;
            double v = (2 - sigma[IJ] * tau[IJ] * dt * dt - 2 * dt * dt / hx / hx * c[IJ]
                        - 2 * dt * dt / hy / hy * c[IJ]) * w[IJ]
                       + c[IJ] * (dt / hx) * (dt / hx) * (w[IpJ] + w[InJ])
                       + c[IJ] * (dt / hy) * (dt / hy) * (w[IJp] + w[IJn])
                       + (dt * dt / (2.0 * hx)) * (phi[IpJ] - phi[InJ])
                       + (dt * dt / (2.0 * hy)) * (psi[IJp] - psi[IJn])
                       - (1 - (sigma[IJ] + tau[IJ]) * dt / 2) * wold[IJ];
            u[IJ] = v / (1 + (sigma[IJ] + tau[IJ]) / 2 * dt);
        }
    }
    /* phi/psi: v = w (kernel 1) or the freshly computed u (kernel 0). */
    const double *v = kernel == 0 ? u : w;
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < NX + 2; i++) {
        for (i64 j = 0; j < NY + 2; j++) {
            i64 IJ = i * S + j;
            if (i == 0 || i == NX + 1 || j == 0 || j == NY + 1) {
                phiout[IJ] = 0.0;
                psiout[IJ] = 0.0;
                continue;
            }
            i64 IpJ = IJ + S, InJ = IJ - S, IJp = IJ + 1, IJn = IJ - 1;
            phiout[IJ] = (1. - dt * sigma[IJ]) * phi[IJ]
                         + dt * c[IJ] * (tau[IJ] - sigma[IJ]) / 2.0 / hx * (v[IpJ] - v[InJ]);
            psiout[IJ] = (1. - dt * tau[IJ]) * psi[IJ]
                         + dt * c[IJ] * (sigma[IJ] - tau[IJ]) / 2.0 / hy * (v[IJp] - v[IJn]);
        }
    }
}

/* ----------------------------------------------------------------------------------------
 * a5: VJP of one acoustic step, scatter form.
 *   kernel 1 : AcousticOneStepCpu.h:51-125 (outputs zeroed first, AcousticOneStepCpu.cpp:364-368)
 *   kernel 0 : reverse mode of Core.jl:534-548.  phi/psi read the new interior u, whose ring
 *              values are constants (scatter_nd into zeros, Core.jl:541), so the phi/psi
 *              cotangent is first pulled back onto interior u, then the u-part is transposed.
 * Outputs are overwritten (zeroed then accumulated).  gu_scratch: N doubles (kernel 0 only).
 * -------------------------------------------------------------------------------------- */
void or_acoustic_step_vjp(int kernel, const double *grad_u, const double *grad_phiout,
                          const double *grad_psiout, const double *w, const double *wold,
                          const double *phi, const double *psi, const double *sigma,
                          const double *tau, const double *c, double dt, double hx, double hy,
                          i64 NX, i64 NY, double *grad_w, double *grad_wold, double *grad_phi,
                          double *grad_psi, double *grad_c, double *scratch)
{
    const i64 S = NY + 2, N = (NX + 2) * (NY + 2);
    memset(grad_w, 0, N * sizeof(double));
    memset(grad_wold, 0, N * sizeof(double));
    memset(grad_phi, 0, N * sizeof(double));
    memset(grad_psi, 0, N * sizeof(double));
    memset(grad_c, 0, N * sizeof(double));

    const double *gu = grad_u;
    const double *v = w;
    double *gu_tot = NULL, *unew = NULL;
    if (kernel == 0) {
        /* recompute the new u (phi/psi of kernel 0 depend on it) */
        unew = scratch;
        gu_tot = scratch + N;
        double *tphi = scratch + 2 * N, *tpsi = scratch + 3 * N;
        or_acoustic_step(0, w, wold, phi, psi, sigma, tau, c, dt, hx, hy, NX, NY, unew, tphi, tpsi);
        memcpy(gu_tot, grad_u, N * sizeof(double));
        for (i64 i = 1; i < NX + 1; i++)
            for (i64 j = 1; j < NY + 1; j++) {
                i64 IJ = i * S + j;
                double cphi = dt * c[IJ] * (tau[IJ] - sigma[IJ]) / 2.0 / hx * grad_phiout[IJ];
                double cpsi = dt * c[IJ] * (sigma[IJ] - tau[IJ]) / 2.0 / hy * grad_psiout[IJ];
                gu_tot[IJ + S] += cphi;
                gu_tot[IJ - S] -= cphi;
                gu_tot[IJ + 1] += cpsi;
                gu_tot[IJ - 1] -= cpsi;
            }
        /* ring u is a constant: its cotangent does not propagate */
        for (i64 i = 0; i < NX + 2; i++)
            for (i64 j = 0; j < NY + 2; j++)
                if (i == 0 || i == NX + 1 || j == 0 || j == NY + 1) gu_tot[i * S + j] = 0.0;
        gu = gu_tot;
        v = unew;
    }

    for (i64 i = 1; i < NX + 1; i++) {
        for (i64 j = 1; j < NY + 1; j++) {
            i64 IJ = i * S + j;
            i64 IpJ = IJ + S, InJ = IJ - S, IJp = IJ + 1, IJn = IJ - 1;
            double den = (1 + (sigma[IJ] + tau[IJ]) / 2 * dt);
            grad_c[IJ] += ((-2 * dt * dt / hx / hx - 2 * dt * dt / hy / hy) * w[IJ]
                           + (dt / hx) * (dt / hx) * (w[IpJ] + w[InJ])
                           + (dt / hy) * (dt / hy) * (w[IJp] + w[IJn])) * gu[IJ] / den;
            grad_w[IJ] += (2 - sigma[IJ] * tau[IJ] * dt * dt - 2 * dt * dt / hx / hx * c[IJ]
                           - 2 * dt * dt / hy / hy * c[IJ]) * gu[IJ] / den;
            grad_w[IpJ] += c[IJ] * (dt / hx) * (dt / hx) * gu[IJ] / den;
            grad_w[InJ] += c[IJ] * (dt / hx) * (dt / hx) * gu[IJ] / den;
            grad_w[IJp] += c[IJ] * (dt / hy) * (dt / hy) * gu[IJ] / den;
            grad_w[IJn] += c[IJ] * (dt / hy) * (dt / hy) * gu[IJ] / den;
            grad_phi[IpJ] += (dt * dt / (2.0 * hx)) * gu[IJ] / den;
            grad_phi[InJ] += -(dt * dt / (2.0 * hx)) * gu[IJ] / den;
            grad_psi[IJp] += (dt * dt / (2.0 * hy)) * gu[IJ] / den;
            grad_psi[IJn] += -(dt * dt / (2.0 * hy)) * gu[IJ] / den;
            grad_wold[IJ] += -(1 - (sigma[IJ] + tau[IJ]) * dt / 2) * gu[IJ] / den;

            grad_phi[IJ] += (1. - dt * sigma[IJ]) * grad_phiout[IJ];
            grad_c[IJ] += dt * (tau[IJ] - sigma[IJ]) / 2.0 / hx * (v[IpJ] - v[InJ]) * grad_phiout[IJ];
            grad_psi[IJ] += (1. - dt * tau[IJ]) * grad_psiout[IJ];
            grad_c[IJ] += dt * (sigma[IJ] - tau[IJ]) / 2.0 / hy * (v[IJp] - v[IJn]) * grad_psiout[IJ];
            if (kernel != 0) {
                grad_w[IpJ] += dt * c[IJ] * (tau[IJ] - sigma[IJ]) / 2.0 / hx * grad_phiout[IJ];
                grad_w[InJ] += -dt * c[IJ] * (tau[IJ] - sigma[IJ]) / 2.0 / hx * grad_phiout[IJ];
                grad_w[IJp] += dt * c[IJ] * (sigma[IJ] - tau[IJ]) / 2.0 / hy * grad_psiout[IJ];
                grad_w[IJn] += -dt * c[IJ] * (sigma[IJ] - tau[IJ]) / 2.0 / hy * grad_psiout[IJ];
            }
        }
    }
}

/* ----------------------------------------------------------------------------------------
 * a6 + a7: AcousticPropagatorSolver time loop, src/Core.jl:562-620.
 *   levels 0 and 1 are zero (Core.jl:607-612); for n = 2..NSTEP one step from levels n-1,
 *   n-2, then u^n[src_s] += dt^2 * srcv[n-1, s] (0-based row; Core.jl:593, 600-601;
 *   ScatterAddOps.h:1-7 accumulates duplicates sequentially).
 *   C is the squared velocity (Core.jl:564).  src_idx are 0-based flat indices
 *   (srci-1)(NY+2)+(srcj-1) of the 1-based (srci, srcj) pairs (Core.jl:600).
 *   Histories are (NSTEP+1) x N.
 * -------------------------------------------------------------------------------------- */
void or_acoustic_forward(int kernel, i64 NX, i64 NY, i64 NSTEP, double dt, double hx, double hy,
                         const double *sigma, const double *tau, const double *C, i64 nsrc,
                         const i64 *src_idx, const double *srcv, double *u_hist,
                         double *phi_hist, double *psi_hist)
{
    const i64 N = (NX + 2) * (NY + 2);
    memset(u_hist, 0, 2 * N * sizeof(double));
    memset(phi_hist, 0, 2 * N * sizeof(double));
    memset(psi_hist, 0, 2 * N * sizeof(double));
    for (i64 n = 2; n <= NSTEP; n++) {
        double *u = u_hist + n * N;
        or_acoustic_step(kernel, u_hist + (n - 1) * N, u_hist + (n - 2) * N, phi_hist + (n - 1) * N,
                         psi_hist + (n - 1) * N, sigma, tau, C, dt, hx, hy, NX, NY, u,
                         phi_hist + n * N, psi_hist + n * N);
        for (i64 s = 0; s < nsrc; s++) u[src_idx[s]] += srcv[(n - 1) * nsrc + s] * (dt * dt);
    }
}

/* a8: SimulatedObservation!(::AcousticPropagator), Core.jl:726-730: rcvv[n, r] = u^n[idx_r]. */
void or_acoustic_receive(i64 NSTEP, i64 N, const double *u_hist, i64 nrcv, const i64 *rcv_idx,
                         double *out)
{
    for (i64 n = 0; n <= NSTEP; n++)
        for (i64 r = 0; r < nrcv; r++) out[n * nrcv + r] = u_hist[n * N + rcv_idx[r]];
}

/* ----------------------------------------------------------------------------------------
 * Reverse-mode sweep of the time loop above: what TF autodiff of Core.jl:562-620 computes.
 *   gu_hist: cotangent of the u history ((NSTEP+1) x N), e.g. the scatter of the receiver
 *   residual (GatherOps backward).  Output gC: d/d(C) where C = c^2.  Per-step VJPs are the
 *   function above; source injection is the identity on u (ScatterAddOps_backward).
 * -------------------------------------------------------------------------------------- */
void or_acoustic_adjoint(int kernel, i64 NX, i64 NY, i64 NSTEP, double dt, double hx, double hy,
                         const double *sigma, const double *tau, const double *C,
                         const double *u_hist, const double *phi_hist, const double *psi_hist,
                         const double *gu_hist, double *gC)
{
    const i64 N = (NX + 2) * (NY + 2);
    double *buf = (double *)calloc((size_t)N * 16, sizeof(double));
    double *gu[3] = {buf, buf + N, buf + 2 * N}; /* levels n, n-1, n-2 (rotating) */
    double *gphi = buf + 3 * N, *gpsi = buf + 4 * N;
    double *gw = buf + 5 * N, *gwold = buf + 6 * N, *gp = buf + 7 * N, *gs = buf + 8 * N;
    double *gc_step = buf + 9 * N, *scratch = buf + 10 * N;
    memset(gC, 0, N * sizeof(double));
    /* gu[0] holds level NSTEP, gu[1] level NSTEP-1 */
    memcpy(gu[0], gu_hist + NSTEP * N, N * sizeof(double));
    if (NSTEP >= 1) memcpy(gu[1], gu_hist + (NSTEP - 1) * N, N * sizeof(double));
    for (i64 n = NSTEP; n >= 2; n--) {
        /* gu[2] <- cotangent of level n-2 from the data term */
        memcpy(gu[2], gu_hist + (n - 2) * N, N * sizeof(double));
        or_acoustic_step_vjp(kernel, gu[0], gphi, gpsi, u_hist + (n - 1) * N,
                             u_hist + (n - 2) * N, phi_hist + (n - 1) * N, psi_hist + (n - 1) * N,
                             sigma, tau, C, dt, hx, hy, NX, NY, gw, gwold, gp, gs, gc_step, scratch);
        for (i64 k = 0; k < N; k++) {
            gu[1][k] += gw[k];
            gu[2][k] += gwold[k];
            gC[k] += gc_step[k];
        }
        memcpy(gphi, gp, N * sizeof(double));
        memcpy(gpsi, gs, N * sizeof(double));
        double *t = gu[0];
        gu[0] = gu[1];
        gu[1] = gu[2];
        gu[2] = t;
    }
    free(buf);
}

/* ----------------------------------------------------------------------------------------
 * Gather form of or_acoustic_step_vjp for kernel 1, for the multi-threaded CPU baseline: the
 * same terms as the scatter form above (AcousticOneStepCpu.h:76-123), each output cell
 * collecting its own, so rows are independent and run in parallel.  Equal to the scatter form
 * up to summation order (pinned against it in tests/test_oracle_acoustic.py).
 * -------------------------------------------------------------------------------------- */
void or_acoustic_step_vjp_gather(const double *grad_u, const double *grad_phiout,
                                 const double *grad_psiout, const double *w, const double *sigma,
                                 const double *tau, const double *c, double dt, double hx,
                                 double hy, i64 NX, i64 NY, double *grad_w, double *grad_wold,
                                 double *grad_phi, double *grad_psi, double *grad_c)
{
    const i64 S = NY + 2;
#define IN(a, b) ((a) >= 1 && (a) <= NX && (b) >= 1 && (b) <= NY)
#define DEN(K) (1 + (sigma[K] + tau[K]) / 2 * dt)
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < NX + 2; i++) {
        for (i64 j = 0; j < NY + 2; j++) {
            const i64 IJ = i * S + j;
            double gw = 0, gwo = 0, gp = 0, gs = 0, gc = 0;
            if (IN(i, j)) {
                const i64 IpJ = IJ + S, InJ = IJ - S, IJp = IJ + 1, IJn = IJ - 1;
                const double den = DEN(IJ);
                gc += ((-2 * dt * dt / hx / hx - 2 * dt * dt / hy / hy) * w[IJ]
                       + (dt / hx) * (dt / hx) * (w[IpJ] + w[InJ])
                       + (dt / hy) * (dt / hy) * (w[IJp] + w[IJn])) * grad_u[IJ] / den;
                gw += (2 - sigma[IJ] * tau[IJ] * dt * dt - 2 * dt * dt / hx / hx * c[IJ]
                       - 2 * dt * dt / hy / hy * c[IJ]) * grad_u[IJ] / den;
                gwo += -(1 - (sigma[IJ] + tau[IJ]) * dt / 2) * grad_u[IJ] / den;
                gp += (1. - dt * sigma[IJ]) * grad_phiout[IJ];
                gc += dt * (tau[IJ] - sigma[IJ]) / 2.0 / hx * (w[IpJ] - w[InJ]) * grad_phiout[IJ];
                gs += (1. - dt * tau[IJ]) * grad_psiout[IJ];
                gc += dt * (sigma[IJ] - tau[IJ]) / 2.0 / hy * (w[IJp] - w[IJn]) * grad_psiout[IJ];
            }
            /* terms of the neighbours K whose stencil reaches IJ */
            if (IN(i - 1, j)) { /* K = (i-1, j): IJ is its IpJ */
                const i64 K = IJ - S;
                gw += c[K] * (dt / hx) * (dt / hx) * grad_u[K] / DEN(K);
                gp += (dt * dt / (2.0 * hx)) * grad_u[K] / DEN(K);
                gw += dt * c[K] * (tau[K] - sigma[K]) / 2.0 / hx * grad_phiout[K];
            }
            if (IN(i + 1, j)) { /* K = (i+1, j): IJ is its InJ */
                const i64 K = IJ + S;
                gw += c[K] * (dt / hx) * (dt / hx) * grad_u[K] / DEN(K);
                gp += -(dt * dt / (2.0 * hx)) * grad_u[K] / DEN(K);
                gw += -dt * c[K] * (tau[K] - sigma[K]) / 2.0 / hx * grad_phiout[K];
            }
            if (IN(i, j - 1)) { /* K = (i, j-1): IJ is its IJp */
                const i64 K = IJ - 1;
                gw += c[K] * (dt / hy) * (dt / hy) * grad_u[K] / DEN(K);
                gs += (dt * dt / (2.0 * hy)) * grad_u[K] / DEN(K);
                gw += dt * c[K] * (sigma[K] - tau[K]) / 2.0 / hy * grad_psiout[K];
            }
            if (IN(i, j + 1)) { /* K = (i, j+1): IJ is its IJn */
                const i64 K = IJ + 1;
                gw += c[K] * (dt / hy) * (dt / hy) * grad_u[K] / DEN(K);
                gs += -(dt * dt / (2.0 * hy)) * grad_u[K] / DEN(K);
                gw += -dt * c[K] * (sigma[K] - tau[K]) / 2.0 / hy * grad_psiout[K];
            }
            grad_w[IJ] = gw;
            grad_wold[IJ] = gwo;
            grad_phi[IJ] = gp;
            grad_psi[IJ] = gs;
            grad_c[IJ] = gc;
        }
    }
#undef IN
#undef DEN
}

/* The reverse sweep of or_acoustic_adjoint for kernel 1 with the gather-form step VJP and
 * parallel element-wise updates: the multi-threaded CPU baseline.  Returns -1 for kernel 0. */
int or_acoustic_adjoint_par(int kernel, i64 NX, i64 NY, i64 NSTEP, double dt, double hx, double hy,
                            const double *sigma, const double *tau, const double *C,
                            const double *u_hist, const double *gu_hist, double *gC)
{
    if (kernel != 1) return -1;
    const i64 N = (NX + 2) * (NY + 2);
    double *buf = (double *)calloc((size_t)N * 10, sizeof(double));
    if (!buf) return -2;
    double *gu[3] = {buf, buf + N, buf + 2 * N};
    double *gphi = buf + 3 * N, *gpsi = buf + 4 * N;
    double *gw = buf + 5 * N, *gwold = buf + 6 * N, *gp = buf + 7 * N, *gs = buf + 8 * N;
    double *gc_step = buf + 9 * N;
    memset(gC, 0, N * sizeof(double));
    memcpy(gu[0], gu_hist + NSTEP * N, N * sizeof(double));
    if (NSTEP >= 1) memcpy(gu[1], gu_hist + (NSTEP - 1) * N, N * sizeof(double));
    for (i64 n = NSTEP; n >= 2; n--) {
        or_acoustic_step_vjp_gather(gu[0], gphi, gpsi, u_hist + (n - 1) * N, sigma, tau, C, dt, hx,
                                    hy, NX, NY, gw, gwold, gp, gs, gc_step);
        const double *d2 = gu_hist + (n - 2) * N;
#pragma omp parallel for schedule(static)
        for (i64 k = 0; k < N; k++) {
            gu[1][k] += gw[k];
            gu[2][k] = d2[k] + gwold[k];
            gC[k] += gc_step[k];
            gphi[k] = gp[k];
            gpsi[k] = gs[k];
        }
        double *t = gu[0];
        gu[0] = gu[1];
        gu[1] = gu[2];
        gu[2] = t;
    }
    free(buf);
    return 0;
}

/* ========================================================================================
 * ELASTIC
 * ======================================================================================== */

/* ----------------------------------------------------------------------------------------
 * a11: CPML coefficients, src/Core.jl:231-407 (one axis).  coordinate x_k = h*k + origin
 * (MPIElastic.jl:229, 297 add an origin).  right edge origin = (NXg-0.5)*h - thickness.
 * alpha grows into the PML (Core.jl:289).  Outputs arrays of length len: a/b/K for the full
 * grid and the half grid (x_k + h/2).
 * -------------------------------------------------------------------------------------- */
void or_elastic_pml_axis(i64 len, double h, double origin, i64 NXg, i64 npml, double npower,
                         double kmax, double alpha_max, double Rcoef, double vp_ref, double dt,
                         int use_min, int use_max, double *a, double *a_half, double *b,
                         double *b_half, double *K, double *K_half)
{
    const double thick = (double)npml * h;
    const double d0 = -(npower + 1) * vp_ref * log(Rcoef) / (2. * thick);
    const double oleft = thick;
    const double oright = ((double)NXg - 0.5) * h - thick;
    for (i64 k = 0; k < len; k++) {
        double xval = h * (double)k + origin;
        double d = 0, dh = 0, kk = 1, kh = 1, al = 0, alh = 0;
        if (use_min) {
            double ab = oleft - xval;
            if (ab >= 0.0) {
                double an = ab / thick;
                d = d0 * pow(an, npower);
                kk = 1. + (kmax - 1.) * pow(an, npower);
                al = alpha_max * an;
            }
            ab = oleft - (xval + h / 2.);
            if (ab >= 0.0) {
                double an = ab / thick;
                dh = d0 * pow(an, npower);
                kh = 1. + (kmax - 1.) * pow(an, npower);
                alh = alpha_max * an;
            }
        }
        if (use_max) {
            double ab = xval - oright;
            if (ab >= 0) {
                double an = ab / thick;
                d = d0 * pow(an, npower);
                kk = 1. + (kmax - 1.) * pow(an, npower);
                al = alpha_max * an;
            }
            ab = xval + h / 2.0 - oright;
            if (ab >= 0) {
                double an = ab / thick;
                dh = d0 * pow(an, npower);
                kh = 1. + (kmax - 1.) * pow(an, npower);
                alh = alpha_max * an;
            }
        }
        if (al < 0) al = 0;
        if (alh < 0) alh = 0;
        b[k] = exp(-(d / kk + al) * dt);
        b_half[k] = exp(-(dh / kh + alh) * dt);
        a[k] = fabs(d) > 1e-6 ? d * (b[k] - 1.) / (kk * (d + kk * al)) : 0.0;
        a_half[k] = fabs(dh) > 1e-6 ? dh * (b_half[k] - 1.) / (kh * (dh + kh * alh)) : 0.0;
        K[k] = kk;
        K_half[k] = kh;
    }
}

/* Elastic state: fields f[0..4] = vx, vy, sxx, syy, sxy (srctype order, Struct.jl:44-48)
 * and memory variables m[0..7] = mem[1..8] of Core.jl.  Profiles: ax/bx have 2*NX entries
 * (row 0 full grid, row 1 half grid, Core.jl:397-402), ay/by 2*NY. */
typedef struct {
    i64 NX, NY;
    double hx, hy, dt;
    const double *rho, *lam, *mu;
    const double *axf, *axh, *bxf, *bxh, *ayf, *ayh, *byf, *byh;
    /* sub = 1: the one-subdomain form of MPIElastic.jl (a block of (NX-2) x (NY-2) cells at
     * [2, NX-1] x [2, NY-1] inside a 2-cell zero halo): every fw updates the whole block
     * (MPIElastic.jl:483-620, index set ij = kk x kk), parameters are not averaged
     * (:499-502, 537-538, 605-606), and the profile index is i-2 / j-2 (length NX-2). */
    int sub;
} el_ctx;

/* update range of fw1..fw4 (Core.jl:100-107, 132-139, 162-169, 190-197; whole block if sub) */
static void el_range(const el_ctx *e, int fw, i64 *ilo, i64 *ihi, i64 *jlo, i64 *jhi)
{
    if (e->sub) {
        *ilo = 2, *ihi = e->NX - 1, *jlo = 2, *jhi = e->NY - 1;
        return;
    }
    switch (fw) {
    case 1: *ilo = 1, *ihi = e->NX - 1, *jlo = 2, *jhi = e->NY; break;
    case 2: *ilo = 2, *ihi = e->NX, *jlo = 1, *jhi = e->NY - 1; break;
    case 3: *ilo = 2, *ihi = e->NX, *jlo = 2, *jhi = e->NY; break;
    default: *ilo = 1, *ihi = e->NX - 1, *jlo = 1, *jhi = e->NY - 1; break;
    }
}
#define EL_PO(e) ((e)->sub ? 2 : 1) /* profile index offset */

/* fw1: sxx, syy, m1, m2.  Core.jl:96-126.  i in [1, NX-1], j in [2, NY] (0-based). */
static void el_fw1(const el_ctx *e, const double *vx, const double *vy, double *sxx, double *syy,
                   double *m1, double *m2)
{
    const i64 S = e->NY + 2, NX = e->NX, NY = e->NY;
    const i64 N = (NX + 2) * S;
    double *n1 = (double *)calloc(N, sizeof(double)), *n2 = (double *)calloc(N, sizeof(double));
    i64 ilo, ihi, jlo, jhi;
    el_range(e, 1, &ilo, &ihi, &jlo, &jhi);
    const i64 po = EL_PO(e);
#pragma omp parallel for schedule(static)
    for (i64 i = ilo; i <= ihi; i++)
        for (i64 j = jlo; j <= jhi; j++) {
            i64 ij = i * S + j;
            double lb = e->sub ? e->lam[ij] : 0.5 * (e->lam[ij + S] + e->lam[ij]);
            double mb = e->sub ? e->mu[ij] : 0.5 * (e->mu[ij + S] + e->mu[ij]);
            double lm = lb + 2 * mb;
            double dx = (27 * vx[ij + S] - 27 * vx[ij] - vx[ij + 2 * S] + vx[ij - S]) / (24 * e->hx);
            double dy = (27 * vy[ij] - 27 * vy[ij - 1] - vy[ij + 1] + vy[ij - 2]) / (24 * e->hy);
            n1[ij] = e->bxh[i - po] * m1[ij] + e->axh[i - po] * dx;
            n2[ij] = e->byf[j - po] * m2[ij] + e->ayf[j - po] * dy;
            dx = dx + n1[ij];
            dy = dy + n2[ij];
            sxx[ij] += (lm * dx + lb * dy) * e->dt;
            syy[ij] += (lm * dy + lb * dx) * e->dt;
        }
    memcpy(m1, n1, N * sizeof(double));
    memcpy(m2, n2, N * sizeof(double));
    free(n1);
    free(n2);
}

/* fw2: sxy, m3, m4.  Core.jl:128-155.  i in [2, NX], j in [1, NY-1]. */
static void el_fw2(const el_ctx *e, const double *vx, const double *vy, double *sxy, double *m3,
                   double *m4)
{
    const i64 S = e->NY + 2, NX = e->NX, NY = e->NY;
    const i64 N = (NX + 2) * S;
    double *n3 = (double *)calloc(N, sizeof(double)), *n4 = (double *)calloc(N, sizeof(double));
    i64 ilo, ihi, jlo, jhi;
    el_range(e, 2, &ilo, &ihi, &jlo, &jhi);
    const i64 po = EL_PO(e);
#pragma omp parallel for schedule(static)
    for (i64 i = ilo; i <= ihi; i++)
        for (i64 j = jlo; j <= jhi; j++) {
            i64 ij = i * S + j;
            double mb = e->sub ? e->mu[ij] : 0.5 * (e->mu[ij] + e->mu[ij + 1]);
            double dx = (27 * vy[ij] - 27 * vy[ij - S] - vy[ij + S] + vy[ij - 2 * S]) / (24 * e->hx);
            double dy = (27 * vx[ij + 1] - 27 * vx[ij] - vx[ij + 2] + vx[ij - 1]) / (24 * e->hy);
            n3[ij] = e->bxf[i - po] * m3[ij] + e->axf[i - po] * dx;
            n4[ij] = e->byh[j - po] * m4[ij] + e->ayh[j - po] * dy;
            dx = dx + n3[ij];
            dy = dy + n4[ij];
            sxy[ij] += mb * (dx + dy) * e->dt;
        }
    memcpy(m3, n3, N * sizeof(double));
    memcpy(m4, n4, N * sizeof(double));
    free(n3);
    free(n4);
}

/* fw3: vx, m5, m6.  Core.jl:158-182.  i in [2, NX], j in [2, NY]. */
static void el_fw3(const el_ctx *e, const double *sxx, const double *sxy, double *vx, double *m5,
                   double *m6)
{
    const i64 S = e->NY + 2, NX = e->NX, NY = e->NY;
    const i64 N = (NX + 2) * S;
    double *n5 = (double *)calloc(N, sizeof(double)), *n6 = (double *)calloc(N, sizeof(double));
    i64 ilo, ihi, jlo, jhi;
    el_range(e, 3, &ilo, &ihi, &jlo, &jhi);
    const i64 po = EL_PO(e);
#pragma omp parallel for schedule(static)
    for (i64 i = ilo; i <= ihi; i++)
        for (i64 j = jlo; j <= jhi; j++) {
            i64 ij = i * S + j;
            double dx = (27 * sxx[ij] - 27 * sxx[ij - S] - sxx[ij + S] + sxx[ij - 2 * S]) / (24 * e->hx);
            double dy = (27 * sxy[ij] - 27 * sxy[ij - 1] - sxy[ij + 1] + sxy[ij - 2]) / (24 * e->hy);
            n5[ij] = e->bxf[i - po] * m5[ij] + e->axf[i - po] * dx;
            n6[ij] = e->byf[j - po] * m6[ij] + e->ayf[j - po] * dy;
            dx = dx + n5[ij];
            dy = dy + n6[ij];
            vx[ij] += (dx + dy) * e->dt / e->rho[ij];
        }
    memcpy(m5, n5, N * sizeof(double));
    memcpy(m6, n6, N * sizeof(double));
    free(n5);
    free(n6);
}

/* fw4: vy, m7, m8.  Core.jl:186-213.  i in [1, NX-1], j in [1, NY-1]. */
static void el_fw4(const el_ctx *e, const double *sxy, const double *syy, double *vy, double *m7,
                   double *m8)
{
    const i64 S = e->NY + 2, NX = e->NX, NY = e->NY;
    const i64 N = (NX + 2) * S;
    double *n7 = (double *)calloc(N, sizeof(double)), *n8 = (double *)calloc(N, sizeof(double));
    i64 ilo, ihi, jlo, jhi;
    el_range(e, 4, &ilo, &ihi, &jlo, &jhi);
    const i64 po = EL_PO(e);
#pragma omp parallel for schedule(static)
    for (i64 i = ilo; i <= ihi; i++)
        for (i64 j = jlo; j <= jhi; j++) {
            i64 ij = i * S + j;
            double rb = e->sub ? e->rho[ij]
                               : 0.25 * (e->rho[ij] + e->rho[ij + S] + e->rho[ij + S + 1] + e->rho[ij + 1]);
            double dx = (27 * sxy[ij + S] - 27 * sxy[ij] - sxy[ij + 2 * S] + sxy[ij - S]) / (24 * e->hx);
            double dy = (27 * syy[ij + 1] - 27 * syy[ij] - syy[ij + 2] + syy[ij - 1]) / (24 * e->hy);
            n7[ij] = e->bxh[i - po] * m7[ij] + e->axh[i - po] * dx;
            n8[ij] = e->byh[j - po] * m8[ij] + e->ayh[j - po] * dy;
            dx = dx + n7[ij];
            dy = dy + n8[ij];
            vy[ij] += (dx + dy) * e->dt / rb;
        }
    memcpy(m7, n7, N * sizeof(double));
    memcpy(m8, n8, N * sizeof(double));
    free(n7);
    free(n8);
}

static void el_ctx_init(el_ctx *e, i64 NX, i64 NY, double hx, double hy, double dt,
                        const double *rho, const double *lam, const double *mu, const double *ax,
                        const double *bx, const double *ay, const double *by, int sub)
{
    const i64 PX = sub ? NX - 2 : NX, PY = sub ? NY - 2 : NY; /* profile lengths */
    e->sub = sub;
    e->NX = NX;
    e->NY = NY;
    e->hx = hx;
    e->hy = hy;
    e->dt = dt;
    e->rho = rho;
    e->lam = lam;
    e->mu = mu;
    e->axf = ax;
    e->axh = ax + PX;
    e->bxf = bx;
    e->bxh = bx + PX;
    e->ayf = ay;
    e->ayh = ay + PY;
    e->byf = by;
    e->byh = by + PY;
}

/* ----------------------------------------------------------------------------------------
 * a12 + a15: one elastic step in place, src/Core.jl:215-228: fw1, fw2, fw3, fw4, then the
 * CPU AddSource semantics (deps/CustomOps/SourceOps/AddSource.cpp:57-85):
 *   field_type[srci-1, srcj-1] += srcv[s], sequentially, 1-based (srci, srcj).
 * f: 5 fields of N (vx, vy, sxx, syy, sxy).  m: 8 memory variables of N.
 * -------------------------------------------------------------------------------------- */
void or_elastic_step(i64 NX, i64 NY, double hx, double hy, double dt, const double *rho,
                     const double *lam, const double *mu, const double *ax, const double *bx,
                     const double *ay, const double *by, double *f, double *m, i64 nsrc,
                     const i64 *srci, const i64 *srcj, const i64 *srctype, const double *srcv,
                     int sub)
{
    el_ctx e;
    el_ctx_init(&e, NX, NY, hx, hy, dt, rho, lam, mu, ax, bx, ay, by, sub);
    const i64 N = (NX + 2) * (NY + 2), S = NY + 2;
    double *vx = f, *vy = f + N, *sxx = f + 2 * N, *syy = f + 3 * N, *sxy = f + 4 * N;
    el_fw1(&e, vx, vy, sxx, syy, m, m + N);
    el_fw2(&e, vx, vy, sxy, m + 2 * N, m + 3 * N);
    el_fw3(&e, sxx, sxy, vx, m + 4 * N, m + 5 * N);
    el_fw4(&e, sxy, syy, vy, m + 6 * N, m + 7 * N);
    for (i64 s = 0; s < nsrc; s++) {
        i64 t = srctype[s];
        if (t < 0 || t > 4) continue;
        f[t * N + (srci[s] - 1) * S + (srcj[s] - 1)] += srcv[s];
    }
}

/* a14: ElasticPropagatorSolver time loop, Core.jl:31-94 + ElasticSourceAtTimeT :664-667.
 * level 0 zero; for n = 1..NSTEP one step, sources take srcv row n-1 (0-based).
 * hist: 5 x (NSTEP+1) x N (field-major).  mem_hist: 8 x (NSTEP+1) x N (may be NULL). */
void or_elastic_forward(i64 NX, i64 NY, i64 NSTEP, double hx, double hy, double dt,
                        const double *rho, const double *lam, const double *mu, const double *ax,
                        const double *bx, const double *ay, const double *by, i64 nsrc,
                        const i64 *srci, const i64 *srcj, const i64 *srctype, const double *srcv,
                        double *hist, double *mem_hist, int sub)
{
    const i64 N = (NX + 2) * (NY + 2), L = (NSTEP + 1) * N;
    double *f = (double *)calloc(5 * N, sizeof(double));
    double *m = (double *)calloc(8 * N, sizeof(double));
    for (int k = 0; k < 5; k++) memset(hist + k * L, 0, N * sizeof(double));
    if (mem_hist)
        for (int k = 0; k < 8; k++) memset(mem_hist + k * L, 0, N * sizeof(double));
    for (i64 n = 1; n <= NSTEP; n++) {
        or_elastic_step(NX, NY, hx, hy, dt, rho, lam, mu, ax, bx, ay, by, f, m, nsrc, srci, srcj,
                        srctype, srcv + (n - 1) * nsrc, sub);
        for (int k = 0; k < 5; k++) memcpy(hist + k * L + n * N, f + k * N, N * sizeof(double));
        if (mem_hist)
            for (int k = 0; k < 8; k++) memcpy(mem_hist + k * L + n * N, m + k * N, N * sizeof(double));
    }
    free(f);
    free(m);
}

/* a16: GetReceive forward, deps/CustomOps/ReceiveOps/GetReceive.cpp:10-46:
 * out[r*(NSTEP+1) + n] = field_type_r[n][(rcvi-1)(NY+2) + rcvj-1]  (layout (nrcv, NSTEP+1)). */
void or_elastic_receive(i64 NX, i64 NY, i64 NSTEP, const double *hist, i64 nrcv, const i64 *rcvi,
                        const i64 *rcvj, const i64 *rcvtype, double *out)
{
    const i64 N = (NX + 2) * (NY + 2), L = (NSTEP + 1) * N, nt = NSTEP + 1;
    for (i64 r = 0; r < nrcv; r++) {
        i64 t = rcvtype[r];
        if (t < 0 || t > 4) continue;
        i64 idx = (rcvi[r] - 1) * (NY + 2) + rcvj[r] - 1;
        for (i64 n = 0; n < nt; n++) out[nt * r + n] = hist[t * L + n * N + idx];
    }
}

/* ---------------- reverse mode of fw1..fw4 (scatter form) ----------------
 * G: cotangents of the 5 fields (in/out), Gm: of the 8 memory variables (in/out).
 * gr/gl/gm: accumulated cotangents of rho, lambda, mu.  Inputs are the values the forward
 * phase read. */
static void el_fw4_vjp(const el_ctx *e, const double *sxy, const double *syy, const double *m7,
                       const double *m8, double *Gvy, double *Gsxy, double *Gsyy, double *Gm7,
                       double *Gm8, double *gr)
{
    const i64 S = e->NY + 2, NX = e->NX, NY = e->NY, N = (NX + 2) * S;
    double *o7 = (double *)calloc(N, sizeof(double)), *o8 = (double *)calloc(N, sizeof(double));
    const double c1 = 27.0 / (24 * e->hx), c2 = 1.0 / (24 * e->hx);
    const double d1 = 27.0 / (24 * e->hy), d2 = 1.0 / (24 * e->hy);
    i64 ilo, ihi, jlo, jhi;
    el_range(e, 4, &ilo, &ihi, &jlo, &jhi);
    const i64 po = EL_PO(e);
    for (i64 i = ilo; i <= ihi; i++)
        for (i64 j = jlo; j <= jhi; j++) {
            i64 ij = i * S + j;
            double rb = e->sub ? e->rho[ij]
                               : 0.25 * (e->rho[ij] + e->rho[ij + S] + e->rho[ij + S + 1] + e->rho[ij + 1]);
            double dx = (27 * sxy[ij + S] - 27 * sxy[ij] - sxy[ij + 2 * S] + sxy[ij - S]) / (24 * e->hx);
            double dy = (27 * syy[ij + 1] - 27 * syy[ij] - syy[ij + 2] + syy[ij - 1]) / (24 * e->hy);
            double n7 = e->bxh[i - po] * m7[ij] + e->axh[i - po] * dx;
            double n8 = e->byh[j - po] * m8[ij] + e->ayh[j - po] * dy;
            double g = Gvy[ij];
            double sum = dx + n7 + dy + n8;
            double grb = -g * sum * e->dt / (rb * rb);
            if (e->sub) {
                gr[ij] += grb;
            } else {
                gr[ij] += 0.25 * grb;
                gr[ij + S] += 0.25 * grb;
                gr[ij + S + 1] += 0.25 * grb;
                gr[ij + 1] += 0.25 * grb;
            }
            double gs = g * e->dt / rb;               /* cotangent of (dx+n7+dy+n8) */
            double g7 = Gm7[ij] + gs, g8 = Gm8[ij] + gs; /* total cotangent of n7, n8 */
            double gdx = gs + e->axh[i - po] * g7;
            double gdy = gs + e->ayh[j - po] * g8;
            o7[ij] = e->bxh[i - po] * g7;
            o8[ij] = e->byh[j - po] * g8;
            Gsxy[ij + S] += c1 * gdx;
            Gsxy[ij] -= c1 * gdx;
            Gsxy[ij + 2 * S] -= c2 * gdx;
            Gsxy[ij - S] += c2 * gdx;
            Gsyy[ij + 1] += d1 * gdy;
            Gsyy[ij] -= d1 * gdy;
            Gsyy[ij + 2] -= d2 * gdy;
            Gsyy[ij - 1] += d2 * gdy;
        }
    memcpy(Gm7, o7, N * sizeof(double));
    memcpy(Gm8, o8, N * sizeof(double));
    free(o7);
    free(o8);
}

static void el_fw3_vjp(const el_ctx *e, const double *sxx, const double *sxy, const double *m5,
                       const double *m6, double *Gvx, double *Gsxx, double *Gsxy, double *Gm5,
                       double *Gm6, double *gr)
{
    const i64 S = e->NY + 2, NX = e->NX, NY = e->NY, N = (NX + 2) * S;
    double *o5 = (double *)calloc(N, sizeof(double)), *o6 = (double *)calloc(N, sizeof(double));
    const double c1 = 27.0 / (24 * e->hx), c2 = 1.0 / (24 * e->hx);
    const double d1 = 27.0 / (24 * e->hy), d2 = 1.0 / (24 * e->hy);
    i64 ilo, ihi, jlo, jhi;
    el_range(e, 3, &ilo, &ihi, &jlo, &jhi);
    const i64 po = EL_PO(e);
    for (i64 i = ilo; i <= ihi; i++)
        for (i64 j = jlo; j <= jhi; j++) {
            i64 ij = i * S + j;
            double dx = (27 * sxx[ij] - 27 * sxx[ij - S] - sxx[ij + S] + sxx[ij - 2 * S]) / (24 * e->hx);
            double dy = (27 * sxy[ij] - 27 * sxy[ij - 1] - sxy[ij + 1] + sxy[ij - 2]) / (24 * e->hy);
            double n5 = e->bxf[i - po] * m5[ij] + e->axf[i - po] * dx;
            double n6 = e->byf[j - po] * m6[ij] + e->ayf[j - po] * dy;
            double g = Gvx[ij];
            double r = e->rho[ij];
            gr[ij] += -g * (dx + n5 + dy + n6) * e->dt / (r * r);
            double gs = g * e->dt / r;
            double g5 = Gm5[ij] + gs, g6 = Gm6[ij] + gs;
            double gdx = gs + e->axf[i - po] * g5;
            double gdy = gs + e->ayf[j - po] * g6;
            o5[ij] = e->bxf[i - po] * g5;
            o6[ij] = e->byf[j - po] * g6;
            Gsxx[ij] += c1 * gdx;
            Gsxx[ij - S] -= c1 * gdx;
            Gsxx[ij + S] -= c2 * gdx;
            Gsxx[ij - 2 * S] += c2 * gdx;
            Gsxy[ij] += d1 * gdy;
            Gsxy[ij - 1] -= d1 * gdy;
            Gsxy[ij + 1] -= d2 * gdy;
            Gsxy[ij - 2] += d2 * gdy;
        }
    memcpy(Gm5, o5, N * sizeof(double));
    memcpy(Gm6, o6, N * sizeof(double));
    free(o5);
    free(o6);
}

static void el_fw2_vjp(const el_ctx *e, const double *vx, const double *vy, const double *m3,
                       const double *m4, double *Gsxy, double *Gvx, double *Gvy, double *Gm3,
                       double *Gm4, double *gm)
{
    const i64 S = e->NY + 2, NX = e->NX, NY = e->NY, N = (NX + 2) * S;
    double *o3 = (double *)calloc(N, sizeof(double)), *o4 = (double *)calloc(N, sizeof(double));
    const double c1 = 27.0 / (24 * e->hx), c2 = 1.0 / (24 * e->hx);
    const double d1 = 27.0 / (24 * e->hy), d2 = 1.0 / (24 * e->hy);
    i64 ilo, ihi, jlo, jhi;
    el_range(e, 2, &ilo, &ihi, &jlo, &jhi);
    const i64 po = EL_PO(e);
    for (i64 i = ilo; i <= ihi; i++)
        for (i64 j = jlo; j <= jhi; j++) {
            i64 ij = i * S + j;
            double mb = e->sub ? e->mu[ij] : 0.5 * (e->mu[ij] + e->mu[ij + 1]);
            double dx = (27 * vy[ij] - 27 * vy[ij - S] - vy[ij + S] + vy[ij - 2 * S]) / (24 * e->hx);
            double dy = (27 * vx[ij + 1] - 27 * vx[ij] - vx[ij + 2] + vx[ij - 1]) / (24 * e->hy);
            double n3 = e->bxf[i - po] * m3[ij] + e->axf[i - po] * dx;
            double n4 = e->byh[j - po] * m4[ij] + e->ayh[j - po] * dy;
            double g = Gsxy[ij];
            double gmb = g * (dx + n3 + dy + n4) * e->dt;
            if (e->sub) {
                gm[ij] += gmb;
            } else {
                gm[ij] += 0.5 * gmb;
                gm[ij + 1] += 0.5 * gmb;
            }
            double gs = g * mb * e->dt;
            double g3 = Gm3[ij] + gs, g4 = Gm4[ij] + gs;
            double gdx = gs + e->axf[i - po] * g3;
            double gdy = gs + e->ayh[j - po] * g4;
            o3[ij] = e->bxf[i - po] * g3;
            o4[ij] = e->byh[j - po] * g4;
            Gvy[ij] += c1 * gdx;
            Gvy[ij - S] -= c1 * gdx;
            Gvy[ij + S] -= c2 * gdx;
            Gvy[ij - 2 * S] += c2 * gdx;
            Gvx[ij + 1] += d1 * gdy;
            Gvx[ij] -= d1 * gdy;
            Gvx[ij + 2] -= d2 * gdy;
            Gvx[ij - 1] += d2 * gdy;
        }
    memcpy(Gm3, o3, N * sizeof(double));
    memcpy(Gm4, o4, N * sizeof(double));
    free(o3);
    free(o4);
}

static void el_fw1_vjp(const el_ctx *e, const double *vx, const double *vy, const double *m1,
                       const double *m2, double *Gsxx, double *Gsyy, double *Gvx, double *Gvy,
                       double *Gm1, double *Gm2, double *gl, double *gm)
{
    const i64 S = e->NY + 2, NX = e->NX, NY = e->NY, N = (NX + 2) * S;
    double *o1 = (double *)calloc(N, sizeof(double)), *o2 = (double *)calloc(N, sizeof(double));
    const double c1 = 27.0 / (24 * e->hx), c2 = 1.0 / (24 * e->hx);
    const double d1 = 27.0 / (24 * e->hy), d2 = 1.0 / (24 * e->hy);
    i64 ilo, ihi, jlo, jhi;
    el_range(e, 1, &ilo, &ihi, &jlo, &jhi);
    const i64 po = EL_PO(e);
    for (i64 i = ilo; i <= ihi; i++)
        for (i64 j = jlo; j <= jhi; j++) {
            i64 ij = i * S + j;
            double lb = e->sub ? e->lam[ij] : 0.5 * (e->lam[ij + S] + e->lam[ij]);
            double mb = e->sub ? e->mu[ij] : 0.5 * (e->mu[ij + S] + e->mu[ij]);
            double lm = lb + 2 * mb;
            double dx = (27 * vx[ij + S] - 27 * vx[ij] - vx[ij + 2 * S] + vx[ij - S]) / (24 * e->hx);
            double dy = (27 * vy[ij] - 27 * vy[ij - 1] - vy[ij + 1] + vy[ij - 2]) / (24 * e->hy);
            double n1 = e->bxh[i - po] * m1[ij] + e->axh[i - po] * dx;
            double n2 = e->byf[j - po] * m2[ij] + e->ayf[j - po] * dy;
            double X = dx + n1, Y = dy + n2;
            double gx = Gsxx[ij] * e->dt, gy = Gsyy[ij] * e->dt;
            double glm = X * gx + Y * gy;
            double glb = Y * gx + X * gy + glm;
            double gmb = 2 * glm;
            if (e->sub) {
                gl[ij] += glb;
                gm[ij] += gmb;
            } else {
                gl[ij] += 0.5 * glb;
                gl[ij + S] += 0.5 * glb;
                gm[ij] += 0.5 * gmb;
                gm[ij + S] += 0.5 * gmb;
            }
            double gX = lm * gx + lb * gy, gY = lb * gx + lm * gy;
            double g1 = Gm1[ij] + gX, g2 = Gm2[ij] + gY;
            double gdx = gX + e->axh[i - po] * g1;
            double gdy = gY + e->ayf[j - po] * g2;
            o1[ij] = e->bxh[i - po] * g1;
            o2[ij] = e->byf[j - po] * g2;
            Gvx[ij + S] += c1 * gdx;
            Gvx[ij] -= c1 * gdx;
            Gvx[ij + 2 * S] -= c2 * gdx;
            Gvx[ij - S] += c2 * gdx;
            Gvy[ij] += d1 * gdy;
            Gvy[ij - 1] -= d1 * gdy;
            Gvy[ij + 1] -= d2 * gdy;
            Gvy[ij - 2] += d2 * gdy;
        }
    memcpy(Gm1, o1, N * sizeof(double));
    memcpy(Gm2, o2, N * sizeof(double));
    free(o1);
    free(o2);
}

/* ----------------------------------------------------------------------------------------
 * Reverse-mode sweep of the elastic time loop (what TF autodiff of Core.jl:31-228 gives).
 * g_hist: cotangents of the 5 field histories (5 x (NSTEP+1) x N), e.g. GetReceive's
 * backward (GetReceive.cpp:48-96).  hist/mem_hist from or_elastic_forward.
 * Outputs grho, glam, gmu (overwritten).  Optional gsrcv (NSTEP x nsrc): AddSource backward,
 * AddSource.cpp:129-155 (grad_srcv[s] = cotangent of the field at the source cell).
 * -------------------------------------------------------------------------------------- */
void or_elastic_adjoint(i64 NX, i64 NY, i64 NSTEP, double hx, double hy, double dt,
                        const double *rho, const double *lam, const double *mu, const double *ax,
                        const double *bx, const double *ay, const double *by, const double *hist,
                        const double *mem_hist, const double *g_hist, double *grho, double *glam,
                        double *gmu, i64 nsrc, const i64 *srci, const i64 *srcj,
                        const i64 *srctype, double *gsrcv, int sub)
{
    el_ctx e;
    el_ctx_init(&e, NX, NY, hx, hy, dt, rho, lam, mu, ax, bx, ay, by, sub);
    const i64 N = (NX + 2) * (NY + 2), L = (NSTEP + 1) * N, S = NY + 2;
    double *G = (double *)calloc(5 * N, sizeof(double));
    double *Gm = (double *)calloc(8 * N, sizeof(double));
    double *f = (double *)malloc(5 * N * sizeof(double));
    double *mm = (double *)malloc(8 * N * sizeof(double));
    double *m12 = (double *)malloc(2 * N * sizeof(double));
    double *m34 = (double *)malloc(2 * N * sizeof(double));
    memset(grho, 0, N * sizeof(double));
    memset(glam, 0, N * sizeof(double));
    memset(gmu, 0, N * sizeof(double));
    for (i64 n = NSTEP; n >= 1; n--) {
        for (int k = 0; k < 5; k++)
            for (i64 q = 0; q < N; q++) G[k * N + q] += g_hist[k * L + n * N + q];
        if (gsrcv)
            for (i64 s = 0; s < nsrc; s++) {
                i64 t = srctype[s];
                gsrcv[(n - 1) * nsrc + s] =
                    (t >= 0 && t <= 4) ? G[t * N + (srci[s] - 1) * S + (srcj[s] - 1)] : 0.0;
            }
        /* rebuild the phase inputs from level n-1 */
        for (int k = 0; k < 5; k++) memcpy(f + k * N, hist + k * L + (n - 1) * N, N * sizeof(double));
        for (int k = 0; k < 8; k++) memcpy(mm + k * N, mem_hist + k * L + (n - 1) * N, N * sizeof(double));
        double *vx = f, *vy = f + N, *sxx = f + 2 * N, *syy = f + 3 * N, *sxy = f + 4 * N;
        memcpy(m12, mm, 2 * N * sizeof(double));
        memcpy(m34, mm + 2 * N, 2 * N * sizeof(double));
        double *m1n = (double *)malloc(2 * N * sizeof(double));
        memcpy(m1n, mm, 2 * N * sizeof(double));
        el_fw1(&e, vx, vy, sxx, syy, m1n, m1n + N); /* sxx, syy now post-fw1 */
        double *m3n = (double *)malloc(2 * N * sizeof(double));
        memcpy(m3n, mm + 2 * N, 2 * N * sizeof(double));
        el_fw2(&e, vx, vy, sxy, m3n, m3n + N);      /* sxy post-fw2 */
        free(m1n);
        free(m3n);
        /* fw4 then fw3 (they read post-stress sigma and pre-step memory 5..8) */
        el_fw4_vjp(&e, sxy, syy, mm + 6 * N, mm + 7 * N, G + N, G + 4 * N, G + 3 * N, Gm + 6 * N,
                   Gm + 7 * N, grho);
        el_fw3_vjp(&e, sxx, sxy, mm + 4 * N, mm + 5 * N, G, G + 2 * N, G + 4 * N, Gm + 4 * N,
                   Gm + 5 * N, grho);
        el_fw2_vjp(&e, vx, vy, m34, m34 + N, G + 4 * N, G, G + N, Gm + 2 * N, Gm + 3 * N, gmu);
        el_fw1_vjp(&e, vx, vy, m12, m12 + N, G + 2 * N, G + 3 * N, G, G + N, Gm, Gm + N, glam, gmu);
    }
    free(G);
    free(Gm);
    free(f);
    free(mm);
    free(m12);
    free(m34);
}

/* ----------------------------------------------------------------------------------------
 * a17: Ricker(param, a, shift, amp), src/Utils.jl:188-204 (sample domain, k = 1..NT).
 * a18: frequency-domain Ricker of the shipped fixtures (examples/nn_fwi/Marmousi_model.ipynb
 *      cell 4): r(t) = (1 - 2 pi^2 f^2 t^2) exp(-pi^2 f^2 t^2) * amp, t in [-T, T) step dt,
 *      T = sqrt(20)/(pi f), zero-padded to nt.
 * -------------------------------------------------------------------------------------- */
void or_ricker(i64 NT, double a, double shift, double amp, double *out)
{
    double A = 2 / (sqrt(3 * a) * pow(M_PI, 0.25));
    double wsq = a * a;
    for (i64 k = 1; k <= NT; k++) {
        double v = (double)k - shift;
        double xsq = v * v;
        out[k - 1] = amp * A * (1 - xsq / wsq) * exp(-xsq / (2 * wsq));
    }
}

i64 or_ricker_freq(i64 nt, double f0, double dt, double amp, double *out)
{
    /* t = arange(-T, T, dt): ceil(2T/dt) samples -T + k*dt, copied into vec[:len] */
    double T = sqrt(20.0) / (M_PI * f0);
    i64 m = (i64)ceil((T - (-T)) / dt);
    memset(out, 0, nt * sizeof(double));
    for (i64 k = 0; k < m && k < nt; k++) {
        double t = -T + (double)k * dt;
        out[k] = (1 - 2 * M_PI * M_PI * f0 * f0 * t * t) * exp(-M_PI * M_PI * f0 * f0 * t * t) * amp;
    }
    return m;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
