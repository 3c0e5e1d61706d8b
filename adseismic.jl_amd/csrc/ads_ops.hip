// ads_ops.hip -- drop-in per-step operators that replace the reference's TensorFlow custom
// ops one for one (deps/CustomOps/*).  This translation unit is compiled with
// -ffp-contract=off and evaluates every expression in the reference's order, so the FP64
// results equal the reference CPU kernels bit for bit.
#include <algorithm>
#include <mutex>
#include <vector>

#include "ads_common.h"

namespace ads {

// AcousticOneStepCpu.h:18-46, one thread per cell.
template <typename T>
__global__ __launch_bounds__(256) void op_ac_step(const T *__restrict__ w, const T *__restrict__ wold,
                                                  const T *__restrict__ phi, const T *__restrict__ psi,
                                                  const T *__restrict__ sigma, const T *__restrict__ tau,
                                                  const T *__restrict__ c, T dt, T hx, T hy, int64_t NX,
                                                  int64_t NY, T *__restrict__ u, T *__restrict__ phiout,
                                                  T *__restrict__ psiout)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= NY + 2 || i >= NX + 2) return;
    const int64_t S = NY + 2, IJ = i * S + j;
    if (i == 0 || i == NX + 1 || j == 0 || j == NY + 1) {
        u[IJ] = T(0);
        phiout[IJ] = T(0);
        psiout[IJ] = T(0);
        return;
    }
    const int64_t IpJ = IJ + S, InJ = IJ - S, IJp = IJ + 1, IJn = IJ - 1;
    T v = (T(2) - sigma[IJ] * tau[IJ] * dt * dt - T(2) * dt * dt / hx / hx * c[IJ] -
           T(2) * dt * dt / hy / hy * c[IJ]) * w[IJ] +
          c[IJ] * (dt / hx) * (dt / hx) * (w[IpJ] + w[InJ]) +
          c[IJ] * (dt / hy) * (dt / hy) * (w[IJp] + w[IJn]) +
          (dt * dt / (T(2) * hx)) * (phi[IpJ] - phi[InJ]) +
          (dt * dt / (T(2) * hy)) * (psi[IJp] - psi[IJn]) -
          (T(1) - (sigma[IJ] + tau[IJ]) * dt / T(2)) * wold[IJ];
    u[IJ] = v / (T(1) + (sigma[IJ] + tau[IJ]) / T(2) * dt);
    phiout[IJ] = (T(1) - dt * sigma[IJ]) * phi[IJ] +
                 dt * c[IJ] * (tau[IJ] - sigma[IJ]) / T(2) / hx * (w[IpJ] - w[InJ]);
    psiout[IJ] = (T(1) - dt * tau[IJ]) * psi[IJ] +
                 dt * c[IJ] * (sigma[IJ] - tau[IJ]) / T(2) / hy * (w[IJp] - w[IJn]);
}

// Gather form of AcousticOneStepCpuBackward (AcousticOneStepCpu.h:76-123).  Each output cell
// sums the contributions the CPU scatter loop would add to it, in the loop's order (cells in
// row-major order; within a cell, the statement order), starting from 0.
template <typename T> struct GradCtx {
    const T *gu, *gphi, *gpsi, *w, *sigma, *tau, *c;
    T dt, hx, hy;
    int64_t NX, NY, S;
    __device__ bool interior(int64_t i, int64_t j) const
    {
        return i >= 1 && i <= NX && j >= 1 && j <= NY;
    }
    __device__ T den(int64_t m) const { return T(1) + (sigma[m] + tau[m]) / T(2) * dt; }
};

template <typename T>
__global__ __launch_bounds__(256) void op_ac_step_grad(GradCtx<T> g, const T *__restrict__ wold,
                                                       T *__restrict__ grad_w, T *__restrict__ grad_wold,
                                                       T *__restrict__ grad_phi, T *__restrict__ grad_psi,
                                                       T *__restrict__ grad_c)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= g.NY + 2 || i >= g.NX + 2) return;
    const int64_t S = g.S, k = i * S + j;
    const T dt = g.dt, hx = g.hx, hy = g.hy;
    const T *c = g.c, *gu = g.gu, *sigma = g.sigma, *tau = g.tau;
    T aw = T(0), aphi = T(0), apsi = T(0), awold = T(0), ac = T(0);
    // m = (i-1, j): k is its IpJ
    if (g.interior(i - 1, j)) {
        const int64_t m = k - S;
        aw += c[m] * (dt / hx) * (dt / hx) * gu[m] / g.den(m);
        aphi += (dt * dt / (T(2) * hx)) * gu[m] / g.den(m);
        aw += dt * c[m] * (tau[m] - sigma[m]) / T(2) / hx * g.gphi[m];
    }
    // m = (i, j-1): k is its IJp
    if (g.interior(i, j - 1)) {
        const int64_t m = k - 1;
        aw += c[m] * (dt / hy) * (dt / hy) * gu[m] / g.den(m);
        apsi += (dt * dt / (T(2) * hy)) * gu[m] / g.den(m);
        aw += dt * c[m] * (sigma[m] - tau[m]) / T(2) / hy * g.gpsi[m];
    }
    // m = k itself
    if (g.interior(i, j)) {
        const int64_t m = k;
        const T *w = g.w;
        const int64_t IpJ = m + S, InJ = m - S, IJp = m + 1, IJn = m - 1;
        ac += ((-T(2) * dt * dt / hx / hx - T(2) * dt * dt / hy / hy) * w[m] +
               (dt / hx) * (dt / hx) * (w[IpJ] + w[InJ]) + (dt / hy) * (dt / hy) * (w[IJp] + w[IJn])) *
              gu[m] / g.den(m);
        aw += (T(2) - sigma[m] * tau[m] * dt * dt - T(2) * dt * dt / hx / hx * c[m] -
               T(2) * dt * dt / hy / hy * c[m]) * gu[m] / g.den(m);
        awold += -(T(1) - (sigma[m] + tau[m]) * dt / T(2)) * gu[m] / g.den(m);
        aphi += (T(1) - dt * sigma[m]) * g.gphi[m];
        ac += dt * (tau[m] - sigma[m]) / T(2) / hx * (w[IpJ] - w[InJ]) * g.gphi[m];
        apsi += (T(1) - dt * tau[m]) * g.gpsi[m];
        ac += dt * (sigma[m] - tau[m]) / T(2) / hy * (w[IJp] - w[IJn]) * g.gpsi[m];
    }
    // m = (i, j+1): k is its IJn
    if (g.interior(i, j + 1)) {
        const int64_t m = k + 1;
        aw += c[m] * (dt / hy) * (dt / hy) * gu[m] / g.den(m);
        apsi += -(dt * dt / (T(2) * hy)) * gu[m] / g.den(m);
        aw += -dt * c[m] * (sigma[m] - tau[m]) / T(2) / hy * g.gpsi[m];
    }
    // m = (i+1, j): k is its InJ
    if (g.interior(i + 1, j)) {
        const int64_t m = k + S;
        aw += c[m] * (dt / hx) * (dt / hx) * gu[m] / g.den(m);
        aphi += -(dt * dt / (T(2) * hx)) * gu[m] / g.den(m);
        aw += -dt * c[m] * (tau[m] - sigma[m]) / T(2) / hx * g.gphi[m];
    }
    (void)wold;
    grad_w[k] = aw;
    grad_wold[k] = awold;
    grad_phi[k] = aphi;
    grad_psi[k] = apsi;
    grad_c[k] = ac;
}

// MpiAcousticOneStep.h:1-42: halo'd (n+2)^2 inputs, n*n c and outputs.
template <typename T>
__global__ __launch_bounds__(256) void op_mpi_ac_step(const T *__restrict__ w, const T *__restrict__ wold,
                                                      const T *__restrict__ phi, const T *__restrict__ psi,
                                                      const T *__restrict__ sigma, const T *__restrict__ tau,
                                                      const T *__restrict__ c, T dt, T hx, T hy, int64_t NX,
                                                      int64_t NY, T *__restrict__ u, T *__restrict__ phiout,
                                                      T *__restrict__ psiout)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int64_t i = (int64_t)blockIdx.y + 1;
    if (j > NY || i > NX) return;
    const int64_t S = NY + 2, IJ = i * S + j;
    const int64_t IpJ = IJ + S, InJ = IJ - S, IJp = IJ + 1, IJn = IJ - 1;
    const int64_t o = (i - 1) * NY + j - 1;
    const T cij = c[o];
    T v = (T(2) - sigma[IJ] * tau[IJ] * dt * dt - T(2) * dt * dt / hx / hx * cij -
           T(2) * dt * dt / hy / hy * cij) * w[IJ] +
          cij * (dt / hx) * (dt / hx) * (w[IpJ] + w[InJ]) + cij * (dt / hy) * (dt / hy) * (w[IJp] + w[IJn]) +
          (dt * dt / (T(2) * hx)) * (phi[IpJ] - phi[InJ]) + (dt * dt / (T(2) * hy)) * (psi[IJp] - psi[IJn]) -
          (T(1) - (sigma[IJ] + tau[IJ]) * dt / T(2)) * wold[IJ];
    u[o] = v / (T(1) + (sigma[IJ] + tau[IJ]) / T(2) * dt);
    phiout[o] = (T(1) - dt * sigma[IJ]) * phi[IJ] + dt * cij * (tau[IJ] - sigma[IJ]) / T(2) / hx * (w[IpJ] - w[InJ]);
    psiout[o] = (T(1) - dt * tau[IJ]) * psi[IJ] + dt * cij * (sigma[IJ] - tau[IJ]) / T(2) / hy * (w[IJp] - w[IJn]);
}

template <typename T>
__global__ void op_gather(const T *__restrict__ v, int64_t nv, const int64_t *__restrict__ ii,
                          int64_t n, T *__restrict__ out)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int64_t q = ii[p] - 1;
    out[p] = (q >= 0 && q < nv) ? v[q] : T(0);
}

template <typename T>
__global__ void op_copy(const T *__restrict__ a, T *__restrict__ b, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) b[p] = a[p];
}

template <typename T> __global__ void op_zero(T *__restrict__ b, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) b[p] = T(0);
}

// sequential in index order: duplicates accumulate / overwrite exactly as the CPU loops do
template <typename T>
__global__ void op_scatter_seq(T *__restrict__ out, int64_t d, const int64_t *__restrict__ ii,
                               const T *__restrict__ upd, int64_t n, int add)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int64_t p = 0; p < n; p++) {
        int64_t q = ii[p] - 1;
        if (q < 0 || q >= d) continue;
        out[q] = add ? out[q] + upd[p] : upd[p];
    }
}

// AddSource.cpp:57-85 (sources applied in order; fields: 0 vx 1 vy 2 sxx 3 syy 4 sxy)
struct SrcEnt {
    int64_t cell;
    int type;
};
template <typename T>
__global__ void op_add_source(T *vx, T *vy, T *sxx, T *syy, T *sxy, const SrcEnt *__restrict__ e,
                              const T *__restrict__ srcv, int64_t nsrc)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    T *f[5] = {vx, vy, sxx, syy, sxy};
    for (int64_t s = 0; s < nsrc; s++)
        if (e[s].type >= 0 && e[s].type <= 4) f[e[s].type][e[s].cell] += srcv[s];
}

template <typename T>
__global__ void op_add_source_grad(const T *gvx, const T *gvy, const T *gsxx, const T *gsyy,
                                   const T *gsxy, const SrcEnt *__restrict__ e, T *__restrict__ g,
                                   int64_t nsrc)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsrc) return;
    const T *f[5] = {gvx, gvy, gsxx, gsyy, gsxy};
    int t = e[s].type;
    g[s] = (t >= 0 && t <= 4) ? f[t][e[s].cell] : T(0);
}

// GetReceive.cpp:13-45: out[r*nt + k] = field_r[k][cell_r]
template <typename T>
__global__ void op_get_receive(const T *vx, const T *vy, const T *sxx, const T *syy, const T *sxy,
                               int64_t nt, int64_t N, const SrcEnt *__restrict__ e, int64_t nrcv,
                               T *__restrict__ out)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nrcv * nt) return;
    int64_t r = q / nt, k = q % nt;
    const T *f[5] = {vx, vy, sxx, syy, sxy};
    int t = e[r].type;
    if (t < 0 || t > 4) return;  // the CPU op leaves such rows untouched
    out[r * nt + k] = f[t][k * N + e[r].cell];
}

template <typename T>
static int one_step(const T *w, const T *wold, const T *phi, const T *psi, const T *sigma,
                    const T *tau, const T *c, double dt, double hx, double hy, int64_t nx,
                    int64_t ny, T *u, T *phiout, T *psiout, ads_stream_t stream)
{
    clear_error();
    if (!w || !wold || !phi || !psi || !sigma || !tau || !c || !u || !phiout || !psiout)
        ADS_FAIL(ADS_ERR_ARG, "ads_acoustic_one_step: null pointer");
    if (nx < 1 || ny < 1 || nx + 2 > 65535) ADS_FAIL(ADS_ERR_SHAPE, "nx, ny out of range");
    dim3 g(cdiv(ny + 2, 256), nx + 2);
    op_ac_step<T><<<g, 256, 0, S(stream)>>>(w, wold, phi, psi, sigma, tau, c, (T)dt, (T)hx, (T)hy,
                                            nx, ny, u, phiout, psiout);
    ADS_LAUNCH_CHECK();
    return ADS_OK;
}

template <typename T>
static int one_step_grad(const T *gu, const T *gphi, const T *gpsi, const T *w, const T *wold,
                         const T *phi, const T *psi, const T *sigma, const T *tau, const T *c,
                         double dt, double hx, double hy, int64_t nx, int64_t ny, T *grad_w,
                         T *grad_wold, T *grad_phi, T *grad_psi, T *grad_c, ads_stream_t stream)
{
    clear_error();
    (void)phi;
    (void)psi;
    if (!gu || !gphi || !gpsi || !w || !sigma || !tau || !c || !grad_w || !grad_wold ||
        !grad_phi || !grad_psi || !grad_c)
        ADS_FAIL(ADS_ERR_ARG, "ads_acoustic_one_step_grad: null pointer");
    if (nx < 1 || ny < 1 || nx + 2 > 65535) ADS_FAIL(ADS_ERR_SHAPE, "nx, ny out of range");
    GradCtx<T> g{gu, gphi, gpsi, w, sigma, tau, c, (T)dt, (T)hx, (T)hy, nx, ny, ny + 2};
    dim3 gr(cdiv(ny + 2, 256), nx + 2);
    op_ac_step_grad<T><<<gr, 256, 0, S(stream)>>>(g, wold, grad_w, grad_wold, grad_phi, grad_psi, grad_c);
    ADS_LAUNCH_CHECK();
    return ADS_OK;
}

template <typename T>
static int mpi_step(const T *w, const T *wold, const T *phi, const T *psi, const T *sigma,
                    const T *tau, const T *c, double dt, double hx, double hy, int64_t n, T *u,
                    T *phiout, T *psiout, ads_stream_t stream)
{
    clear_error();
    if (!w || !wold || !phi || !psi || !sigma || !tau || !c || !u || !phiout || !psiout)
        ADS_FAIL(ADS_ERR_ARG, "ads_mpi_acoustic_one_step: null pointer");
    if (n < 1 || n > 65535) ADS_FAIL(ADS_ERR_SHAPE, "n out of range");
    dim3 g(cdiv(n, 256), n);
    op_mpi_ac_step<T><<<g, 256, 0, S(stream)>>>(w, wold, phi, psi, sigma, tau, c, (T)dt, (T)hx,
                                                (T)hy, n, n, u, phiout, psiout);
    ADS_LAUNCH_CHECK();
    return ADS_OK;
}

template <typename T>
static int gather(const T *v, int64_t nv, const int64_t *ii, int64_t n, T *out, ads_stream_t st)
{
    clear_error();
    if (n == 0) return ADS_OK;
    if (!v || !ii || !out || n < 0) ADS_FAIL(ADS_ERR_ARG, "ads_gather: bad argument");
    op_gather<T><<<cdiv(n, 256), 256, 0, S(st)>>>(v, nv, ii, n, out);
    ADS_LAUNCH_CHECK();
    return ADS_OK;
}

template <typename T>
static int scatter_add(const T *ipt, int64_t d, const int64_t *ii, const T *upd, int64_t n, T *out,
                       ads_stream_t st)
{
    clear_error();
    if (!ipt || !out || d < 0 || n < 0 || (n > 0 && (!ii || !upd)))
        ADS_FAIL(ADS_ERR_ARG, "ads_scatter_add: bad argument");
    if (d > 0) op_copy<T><<<cdiv(d, 256), 256, 0, S(st)>>>(ipt, out, d);
    if (n > 0) op_scatter_seq<T><<<1, 1, 0, S(st)>>>(out, d, ii, upd, n, 1);
    ADS_LAUNCH_CHECK();
    return ADS_OK;
}

template <typename T>
static int scatter_nd(const int64_t *ii, const T *upd, int64_t n, int64_t m, T *out, ads_stream_t st)
{
    clear_error();
    if (!out || m < 0 || n < 0 || (n > 0 && (!ii || !upd)))
        ADS_FAIL(ADS_ERR_ARG, "ads_scatter_nd: bad argument");
    if (m > 0) op_zero<T><<<cdiv(m, 256), 256, 0, S(st)>>>(out, m);
    if (n > 0) op_scatter_seq<T><<<1, 1, 0, S(st)>>>(out, m, ii, upd, n, 0);
    ADS_LAUNCH_CHECK();
    return ADS_OK;
}

// Host-side validation of 1-based (i, j, type) triples into a small device table owned by a
// per-thread cache (grown, never shrunk).  The kernel that reads the table may still run when
// the same thread calls again on another stream, so release_entries() records an event after
// that kernel and the next upload waits for it before overwriting or reallocating the table.
struct EntryCache {
    DevBuf buf;
    hipEvent_t ev = nullptr;
    bool pending = false;  // the event (kept until the process ends) guards the table
};
static thread_local EntryCache entry_cache;

static int release_entries(hipStream_t st)
{
    EntryCache &c = entry_cache;
    if (!c.ev) ADS_HIP(hipEventCreateWithFlags(&c.ev, hipEventDisableTiming));
    ADS_HIP(hipEventRecord(c.ev, st));
    c.pending = true;
    return ADS_OK;
}

static int upload_entries(const int64_t *I, const int64_t *J, const int64_t *type, int64_t n,
                          int64_t nx, int64_t ny, const char *what, SrcEnt **dev, hipStream_t st)
{
    EntryCache &c = entry_cache;
    DevBuf &buf = c.buf;
    std::vector<SrcEnt> h(n);
    for (int64_t s = 0; s < n; s++) {
        if (I[s] < 1 || I[s] > nx + 2 || J[s] < 1 || J[s] > ny + 2)
            ADS_FAIL(ADS_ERR_INDEX, "%s %lld at (%lld, %lld) outside the padded grid", what,
                     (long long)s, (long long)I[s], (long long)J[s]);
        h[s].cell = (I[s] - 1) * (ny + 2) + (J[s] - 1);
        h[s].type = (int)type[s];
    }
    if (c.pending) {  // the previous call's kernel may still read the table
        ADS_HIP(hipEventSynchronize(c.ev));
        c.pending = false;
    }
    if (buf.bytes < n * sizeof(SrcEnt)) ADS_TRY(buf.alloc(std::max<size_t>(n, 64) * sizeof(SrcEnt)));
    ADS_HIP(hipMemcpyAsync(buf.p, h.data(), n * sizeof(SrcEnt), hipMemcpyHostToDevice, st));
    ADS_HIP(hipStreamSynchronize(st));  // h is a stack copy
    *dev = buf.as<SrcEnt>();
    return ADS_OK;
}

// ---- fixed source table of the allocation-free AddSource entry point ----
// One table per device, in the code object's static data: each device holds its own copy
// until the process ends.  `scratch` is only a host-side stand-in for it; host code uses the
// device address hipGetSymbolAddress returns, fetched once per device with that device current.
constexpr int64_t NOALLOC_MAX_SOURCES = ADS_NOALLOC_MAX_SOURCES;  // 1 MiB of 16-byte entries
constexpr int NOALLOC_MAX_DEVICES = 64;
static_assert(sizeof(SrcEnt) == 16, "table entries are 16 bytes");
static __device__ __align__(256) SrcEnt scratch[NOALLOC_MAX_SOURCES];
static std::mutex scratch_mu;
static SrcEnt *scratch_dev[NOALLOC_MAX_DEVICES] = {};

static int scratch_table(SrcEnt **out)
{
    int dev = -1;
    ADS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        ADS_FAIL(ADS_ERR_ARG, "ads_add_source_noalloc: device %d beyond the %d supported", dev,
                 NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        ADS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_dev[dev] = static_cast<SrcEnt *>(p);
    }
    *out = scratch_dev[dev];
    return ADS_OK;
}

// Validated (cell, type) entries into the device table `dev` through a stack staging block;
// nothing is allocated on the host or the device.
static int upload_entries_fixed(const int64_t *I, const int64_t *J, const int64_t *type,
                                int64_t n, int64_t nx, int64_t ny, SrcEnt *dev, hipStream_t st)
{
    constexpr int64_t CH = 256;
    SrcEnt h[CH];
    for (int64_t s0 = 0; s0 < n; s0 += CH) {
        const int64_t m = std::min(CH, n - s0);
        for (int64_t k = 0; k < m; k++) {
            h[k].cell = (I[s0 + k] - 1) * (ny + 2) + (J[s0 + k] - 1);
            h[k].type = (int)type[s0 + k];
        }
        ADS_HIP(hipMemcpyAsync(dev + s0, h, m * sizeof(SrcEnt), hipMemcpyHostToDevice, st));
        ADS_HIP(hipStreamSynchronize(st));  // h is reused by the next block
    }
    return ADS_OK;
}

}  // namespace ads

using namespace ads;

extern "C" {

int ads_acoustic_one_step_f64(const double *w, const double *wold, const double *phi,
                              const double *psi, const double *sigma, const double *tau,
                              const double *c, double dt, double hx, double hy, int64_t nx,
                              int64_t ny, double *u, double *phiout, double *psiout,
                              ads_stream_t stream)
{
    return one_step<double>(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, nx, ny, u, phiout, psiout, stream);
}
int ads_acoustic_one_step_f32(const float *w, const float *wold, const float *phi, const float *psi,
                              const float *sigma, const float *tau, const float *c, double dt,
                              double hx, double hy, int64_t nx, int64_t ny, float *u, float *phiout,
                              float *psiout, ads_stream_t stream)
{
    return one_step<float>(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, nx, ny, u, phiout, psiout, stream);
}

int ads_acoustic_one_step_grad_f64(const double *gu, const double *gphi, const double *gpsi,
                                   const double *w, const double *wold, const double *phi,
                                   const double *psi, const double *sigma, const double *tau,
                                   const double *c, double dt, double hx, double hy, int64_t nx,
                                   int64_t ny, double *grad_w, double *grad_wold, double *grad_phi,
                                   double *grad_psi, double *grad_c, ads_stream_t stream)
{
    return one_step_grad<double>(gu, gphi, gpsi, w, wold, phi, psi, sigma, tau, c, dt, hx, hy, nx,
                                 ny, grad_w, grad_wold, grad_phi, grad_psi, grad_c, stream);
}
int ads_acoustic_one_step_grad_f32(const float *gu, const float *gphi, const float *gpsi,
                                   const float *w, const float *wold, const float *phi,
                                   const float *psi, const float *sigma, const float *tau,
                                   const float *c, double dt, double hx, double hy, int64_t nx,
                                   int64_t ny, float *grad_w, float *grad_wold, float *grad_phi,
                                   float *grad_psi, float *grad_c, ads_stream_t stream)
{
    return one_step_grad<float>(gu, gphi, gpsi, w, wold, phi, psi, sigma, tau, c, dt, hx, hy, nx,
                                ny, grad_w, grad_wold, grad_phi, grad_psi, grad_c, stream);
}

int ads_mpi_acoustic_one_step_f64(const double *w, const double *wold, const double *phi,
                                  const double *psi, const double *sigma, const double *tau,
                                  const double *c, double dt, double hx, double hy, int64_t n,
                                  double *u, double *phiout, double *psiout, ads_stream_t stream)
{
    return mpi_step<double>(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, n, u, phiout, psiout, stream);
}
int ads_mpi_acoustic_one_step_f32(const float *w, const float *wold, const float *phi,
                                  const float *psi, const float *sigma, const float *tau,
                                  const float *c, double dt, double hx, double hy, int64_t n,
                                  float *u, float *phiout, float *psiout, ads_stream_t stream)
{
    return mpi_step<float>(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, n, u, phiout, psiout, stream);
}

int ads_gather_f64(const double *v, int64_t nv, const int64_t *ii, int64_t n, double *out,
                   ads_stream_t stream)
{
    return gather<double>(v, nv, ii, n, out, stream);
}
int ads_gather_f32(const float *v, int64_t nv, const int64_t *ii, int64_t n, float *out,
                   ads_stream_t stream)
{
    return gather<float>(v, nv, ii, n, out, stream);
}
int ads_scatter_add_f64(const double *ipt, int64_t d, const int64_t *ii, const double *upd,
                        int64_t n, double *out, ads_stream_t stream)
{
    return scatter_add<double>(ipt, d, ii, upd, n, out, stream);
}
int ads_scatter_add_f32(const float *ipt, int64_t d, const int64_t *ii, const float *upd,
                        int64_t n, float *out, ads_stream_t stream)
{
    return scatter_add<float>(ipt, d, ii, upd, n, out, stream);
}
int ads_scatter_nd_f64(const int64_t *ii, const double *upd, int64_t n, int64_t m, double *out,
                       ads_stream_t stream)
{
    return scatter_nd<double>(ii, upd, n, m, out, stream);
}
int ads_scatter_nd_f32(const int64_t *ii, const float *upd, int64_t n, int64_t m, float *out,
                       ads_stream_t stream)
{
    return scatter_nd<float>(ii, upd, n, m, out, stream);
}

int ads_add_source_f64(const double *sxx, const double *syy, const double *sxy, const double *vx,
                       const double *vy, const int64_t *srci, const int64_t *srcj,
                       const int64_t *srctype, int64_t nsrc, int64_t nx, int64_t ny,
                       const double *srcv, double *sxx_out, double *syy_out, double *sxy_out,
                       double *vx_out, double *vy_out, ads_stream_t stream)
{
    clear_error();
    if (!sxx || !syy || !sxy || !vx || !vy || !sxx_out || !syy_out || !sxy_out || !vx_out || !vy_out)
        ADS_FAIL(ADS_ERR_ARG, "ads_add_source: null field pointer");
    if (nsrc > 0 && (!srci || !srcj || !srctype || !srcv))
        ADS_FAIL(ADS_ERR_ARG, "ads_add_source: null source arrays");
    const int64_t N = (nx + 2) * (ny + 2);
    hipStream_t st = S(stream);
    const double *in[5] = {sxx, syy, sxy, vx, vy};
    double *out[5] = {sxx_out, syy_out, sxy_out, vx_out, vy_out};
    for (int q = 0; q < 5; q++)
        if (in[q] != out[q])
            ADS_HIP(hipMemcpyAsync(out[q], in[q], N * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (nsrc <= 0) return ADS_OK;
    SrcEnt *e = nullptr;
    ADS_TRY(upload_entries(srci, srcj, srctype, nsrc, nx, ny, "source", &e, st));
    op_add_source<double><<<1, 1, 0, st>>>(vx_out, vy_out, sxx_out, syy_out, sxy_out, e, srcv, nsrc);
    ADS_LAUNCH_CHECK();
    ADS_TRY(release_entries(st));
    return ADS_OK;
}

int ads_add_source_noalloc_f64(const double *sxx, const double *syy, const double *sxy,
                               const double *vx, const double *vy, const int64_t *srci,
                               const int64_t *srcj, const int64_t *srctype, int64_t nsrc,
                               int64_t nx, int64_t ny, const double *srcv, double *sxx_out,
                               double *syy_out, double *sxy_out, double *vx_out, double *vy_out,
                               ads_stream_t stream)
{
    clear_error();
    if (!sxx || !syy || !sxy || !vx || !vy || !sxx_out || !syy_out || !sxy_out || !vx_out || !vy_out)
        ADS_FAIL(ADS_ERR_ARG, "ads_add_source_noalloc: null field pointer");
    if (nsrc > 0 && (!srci || !srcj || !srctype || !srcv))
        ADS_FAIL(ADS_ERR_ARG, "ads_add_source_noalloc: null source arrays");
    if (nsrc > NOALLOC_MAX_SOURCES)
        ADS_FAIL(ADS_ERR_SHAPE, "ads_add_source_noalloc: %lld sources exceed the fixed table of %lld",
                 (long long)nsrc, (long long)NOALLOC_MAX_SOURCES);
    for (int64_t s = 0; s < nsrc; s++)  // validate everything before touching the outputs
        if (srci[s] < 1 || srci[s] > nx + 2 || srcj[s] < 1 || srcj[s] > ny + 2)
            ADS_FAIL(ADS_ERR_INDEX, "source %lld at (%lld, %lld) outside the padded grid",
                     (long long)s, (long long)srci[s], (long long)srcj[s]);
    const int64_t N = (nx + 2) * (ny + 2);
    hipStream_t st = S(stream);
    const double *in[5] = {sxx, syy, sxy, vx, vy};
    double *out[5] = {sxx_out, syy_out, sxy_out, vx_out, vy_out};
    for (int q = 0; q < 5; q++)
        if (in[q] != out[q])
            ADS_HIP(hipMemcpyAsync(out[q], in[q], N * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (nsrc <= 0) return ADS_OK;
    SrcEnt *e = nullptr;
    ADS_TRY(scratch_table(&e));
    ADS_TRY(upload_entries_fixed(srci, srcj, srctype, nsrc, nx, ny, e, st));
    op_add_source<double><<<1, 1, 0, st>>>(vx_out, vy_out, sxx_out, syy_out, sxy_out, e, srcv, nsrc);
    ADS_LAUNCH_CHECK();
    return ADS_OK;
}

int ads_add_source_grad_f64(const double *g_sxx, const double *g_syy, const double *g_sxy,
                            const double *g_vx, const double *g_vy, const int64_t *srci,
                            const int64_t *srcj, const int64_t *srctype, int64_t nsrc, int64_t nx,
                            int64_t ny, double *grad_srcv, ads_stream_t stream)
{
    clear_error();
    if (nsrc <= 0) return ADS_OK;
    if (!g_sxx || !g_syy || !g_sxy || !g_vx || !g_vy || !srci || !srcj || !srctype || !grad_srcv)
        ADS_FAIL(ADS_ERR_ARG, "ads_add_source_grad: null pointer");
    hipStream_t st = S(stream);
    SrcEnt *e = nullptr;
    ADS_TRY(upload_entries(srci, srcj, srctype, nsrc, nx, ny, "source", &e, st));
    op_add_source_grad<double><<<cdiv(nsrc, 256), 256, 0, st>>>(g_vx, g_vy, g_sxx, g_syy, g_sxy, e,
                                                                grad_srcv, nsrc);
    ADS_LAUNCH_CHECK();
    ADS_TRY(release_entries(st));
    return ADS_OK;
}

int ads_get_receive_f64(const double *vx, const double *vy, const double *sxx, const double *syy,
                        const double *sxy, int64_t nt, const int64_t *rcvi, const int64_t *rcvj,
                        const int64_t *rcvtype, int64_t nrcv, int64_t nx, int64_t ny, double *out,
                        ads_stream_t stream)
{
    clear_error();
    if (nrcv <= 0 || nt <= 0) return ADS_OK;
    if (!vx || !vy || !sxx || !syy || !sxy || !rcvi || !rcvj || !rcvtype || !out)
        ADS_FAIL(ADS_ERR_ARG, "ads_get_receive: null pointer");
    hipStream_t st = S(stream);
    SrcEnt *e = nullptr;
    ADS_TRY(upload_entries(rcvi, rcvj, rcvtype, nrcv, nx, ny, "receiver", &e, st));
    op_get_receive<double><<<cdiv(nrcv * nt, 256), 256, 0, st>>>(vx, vy, sxx, syy, sxy, nt,
                                                                 (nx + 2) * (ny + 2), e, nrcv, out);
    ADS_LAUNCH_CHECK();
    ADS_TRY(release_entries(st));
    return ADS_OK;
}

}  // extern "C"
