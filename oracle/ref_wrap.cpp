// Compiles the reference's header-only CPU kernels exactly as they lie in the reference tree
// (deps/CustomOps/*/ *.h) into oracle/_ref/libadseismic_ref.so, used only to write and check
// the golden vectors of tests/golden/.  The headers use TensorFlow's name `int64` and include
// nothing themselves; the one typedef below gives that name its TensorFlow meaning.
typedef long long int64;

#include "AcousticOneStepCpu.h"    // deps/CustomOps/AcousticOneStepCpu/AcousticOneStepCpu.h:1-125
#include "MpiAcousticOneStep.h"    // deps/CustomOps/MPIAcousticOneStepCpu/MpiAcousticOneStep.h:1-121
#include "GatherOps.h"             // deps/CustomOps/GatherOps/GatherOps.h:1-13
#include "ScatterAddOps.h"         // deps/CustomOps/ScatterAddOps/ScatterAddOps.h:1-19
#include "ScatterNdOps.h"          // deps/CustomOps/ScatterNdOps/ScatterNdOps.h:1-17

extern "C" {
void ref_acoustic_one_step_forward(const double *w, const double *wold, const double *phi,
                                   const double *psi, const double *sigma, const double *tau,
                                   const double *c, double dt, double hx, double hy, int64 NX,
                                   int64 NY, double *u, double *phiout, double *psiout)
{
    AcousticOneStepCpuForward(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, NX, NY, u, phiout, psiout);
}

void ref_acoustic_one_step_backward(double *grad_w, double *grad_wold, double *grad_phi,
                                    double *grad_psi, double *grad_c, const double *grad_u,
                                    const double *grad_phiout, const double *grad_psiout,
                                    const double *w, const double *wold, const double *phi,
                                    const double *psi, const double *sigma, const double *tau,
                                    const double *c, double dt, double hx, double hy, int64 NX,
                                    int64 NY, const double *u, const double *phiout,
                                    const double *psiout)
{
    AcousticOneStepCpuBackward(grad_w, grad_wold, grad_phi, grad_psi, grad_c, grad_u, grad_phiout,
                               grad_psiout, w, wold, phi, psi, sigma, tau, c, dt, hx, hy, NX, NY,
                               u, phiout, psiout);
}

void ref_mpi_acoustic_one_step_forward(const double *w, const double *wold, const double *phi,
                                       const double *psi, const double *sigma, const double *tau,
                                       const double *c, double dt, double hx, double hy, int64 NX,
                                       int64 NY, double *u, double *phiout, double *psiout)
{
    MpiAcousticOneStepCpuForward(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, NX, NY, u, phiout,
                                 psiout);
}

void ref_gather_forward(double *out, const double *v, const int64 *ii, int n)
{
    GatherOps_forward(out, v, ii, n);
}

void ref_scatter_add_forward(double *out, const double *ipt, const int64 *ii, const double *upd,
                             int d, int n)
{
    ScatterAddOps_forward(out, ipt, ii, upd, d, n);
}

void ref_scatter_nd_forward(double *out, const int64 *ii, const double *upd, int n)
{
    ScatterNdOps_forward(out, ii, upd, n);
}
}
