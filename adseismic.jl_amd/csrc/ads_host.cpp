// ads_host.cpp -- host-side pieces of the C-ABI: error state, version, PML/CPML coefficient
// tables and geometry tables.  No device code.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstring>
#include <numeric>

#include "ads_common.h"
#include "ads_schedule.h"

namespace ads {

static thread_local char g_err[1024] = {0};

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = 0; }

// ---- event-based kernel timing ----
namespace {
struct ProfRec {
    int cls;
    hipEvent_t a, b;
    double bytes;
    int levels;  // time levels the timed launches advance
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_recs;
std::vector<hipEvent_t> g_pool;
double g_ms[ADS_PROF_COUNT], g_bytes[ADS_PROF_COUNT];
int64_t g_n[ADS_PROF_COUNT];
int64_t g_calls[ADS_PROF_COUNT];  // launches seen per class since the last begin
int g_every = 1;                  // time launch numbers 0, every, 2 every, ... of each class

hipEvent_t pool_get()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool prof_enabled() { return g_prof_on; }

int prof_start(int cls, hipStream_t st)
{
    if (!g_prof_on || cls < 0 || cls >= ADS_PROF_COUNT) return -1;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (g_calls[cls]++ % g_every != 0) return -1;
    ProfRec r{cls, pool_get(), pool_get(), 0.0, 1};
    if (!r.a || !r.b) return -1;
    if (hipEventRecord(r.a, st) != hipSuccess) return -1;
    g_recs.push_back(r);
    return (int)g_recs.size() - 1;
}

void prof_stop(int rec, hipStream_t st, double bytes, int levels)
{
    if (rec < 0) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (rec >= (int)g_recs.size()) return;
    g_recs[rec].bytes = bytes;
    g_recs[rec].levels = levels;
    (void)hipEventRecord(g_recs[rec].b, st);
}

CellLists build_cell_lists(int64_t nshots, const int64_t *ptr, const std::vector<int64_t> &flat)
{
    CellLists L;
    L.shot_ptr.push_back(0);
    L.ent_ptr.push_back(0);
    for (int64_t s = 0; s < nshots; s++) {
        int64_t b = ptr[s], e = ptr[s + 1];
        std::vector<int64_t> ord(e - b);
        std::iota(ord.begin(), ord.end(), 0);
        // stable sort by cell keeps the original order of entries that share a cell
        std::stable_sort(ord.begin(), ord.end(),
                         [&](int64_t x, int64_t y) { return flat[b + x] < flat[b + y]; });
        for (size_t k = 0; k < ord.size();) {
            int64_t cell = flat[b + ord[k]];
            size_t q = k;
            while (q < ord.size() && flat[b + ord[q]] == cell) {
                L.ent.push_back(ord[q]);
                q++;
            }
            L.cell.push_back(cell);
            L.ent_ptr.push_back((int64_t)L.ent.size());
            k = q;
        }
        L.shot_ptr.push_back((int64_t)L.cell.size());
    }
    return L;
}

}  // namespace ads

using namespace ads;

extern "C" {

const char *ads_last_error(void) { return g_err; }

int ads_version(void) { return 100; /* 0.1.0 */ }

int ads_profile_begin_sampled(int every)
{
    clear_error();
    if (every < 1) ADS_FAIL(ADS_ERR_ARG, "profile sampling stride %d must be >= 1", every);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (int k = 0; k < ADS_PROF_COUNT; k++) g_ms[k] = g_bytes[k] = 0.0, g_n[k] = g_calls[k] = 0;
    g_recs.clear();
    g_every = every;
    g_prof_on = true;
    return ADS_OK;
}

int ads_profile_begin(void) { return ads_profile_begin_sampled(1); }

int ads_profile_end(void)
{
    clear_error();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = false;
    int rc = ADS_OK;
    for (auto &r : g_recs) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) {
            rc = ADS_ERR_HIP;
            set_error("event timing failed");
        } else if (r.cls >= 0 && r.cls < ADS_PROF_COUNT) {
            g_ms[r.cls] += ms;
            g_n[r.cls] += r.levels;
            g_bytes[r.cls] += r.bytes;
        }
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_recs.clear();
    return rc;
}

int ads_profile_query(int cls, double *total_ms, int64_t *launches, double *alg_bytes)
{
    clear_error();
    if (cls < 0 || cls >= ADS_PROF_COUNT) ADS_FAIL(ADS_ERR_ARG, "bad kernel class %d", cls);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (total_ms) *total_ms = g_ms[cls];
    if (launches) *launches = g_n[cls];
    if (alg_bytes) *alg_bytes = g_bytes[cls];
    return ADS_OK;
}

// pml_helper (src/Core.jl:738-761) on one axis; coordinate k*h + origin, far edge at
// h*(NXg+1) - L.
static void acoustic_axis(int64_t len, double h, double origin, int64_t NXg, int64_t npml,
                          double R, double vp_ref, int use_min, int use_max, double *out)
{
    const double L = (double)npml * h;
    const double xi = vp_ref / L * std::log(1.0 / R);  // damping, Core.jl:625-629
    const double right = h * (double)(NXg + 1) - L;
    for (int64_t k = 0; k < len; k++) {
        double x = (double)k * h + origin;
        double v = 0.0;
        if (x < L && use_min) {
            double d = std::fabs(L - x);
            v = xi * (d / L - std::sin(2.0 * M_PI * d / L) / (2.0 * M_PI));
        } else if (x > right && use_max) {
            double d = std::fabs(x - right);
            v = xi * (d / L - std::sin(2.0 * M_PI * d / L) / (2.0 * M_PI));
        }
        out[k] = v;
    }
}

int ads_checkpoint_schedule(int64_t M, int64_t S, int64_t *ops, int64_t cap, int64_t *nops)
{
    clear_error();
    if (M < 0 || !nops || (cap > 0 && !ops)) ADS_FAIL(ADS_ERR_ARG, "bad argument");
    struct Rec {
        int64_t *ops, cap, n = 0;
        void put(int64_t k, int64_t a, int64_t b)
        {
            if (n < cap) {
                ops[3 * n] = k;
                ops[3 * n + 1] = a;
                ops[3 * n + 2] = b;
            }
            n++;
        }
        int restore(int64_t m, int64_t q) { put(0, m, q); return ADS_OK; }
        int forward(int64_t m0, int64_t m1) { put(1, m0, m1); return ADS_OK; }
        int save(int64_t m, int64_t q) { put(2, m, q); return ADS_OK; }
        int adjoint(int64_t m) { put(3, m, 0); return ADS_OK; }
    } rec{ops, cap};
    Reversal rv(M, S < 0 ? std::max<int64_t>(M - 1, 0) : std::min(S, std::max<int64_t>(M - 1, 0)));
    for (const auto &ps : rv.first_path()) rec.put(4, ps.first, ps.second);
    rv.run(rec);
    *nops = rec.n;
    if (rec.n > cap) ADS_FAIL(ADS_ERR_SHAPE, "schedule has %lld operations, room for %lld",
                              (long long)rec.n, (long long)cap);
    return ADS_OK;
}

int ads_acoustic_pml_profiles(const ads_acoustic_params *p, double *sigma_x, double *tau_y)
{
    clear_error();
    if (!p || !sigma_x || !tau_y) ADS_FAIL(ADS_ERR_ARG, "ads_acoustic_pml_profiles: null argument");
    if (p->NX < 1 || p->NY < 1) ADS_FAIL(ADS_ERR_SHAPE, "NX, NY must be >= 1");
    if (p->NPOINTS_PML < 1 || !(p->Rcoef > 0) || !(p->DELTAX > 0) || !(p->DELTAY > 0))
        ADS_FAIL(ADS_ERR_ARG, "bad PML parameters (NPOINTS_PML >= 1, Rcoef > 0, h > 0)");
    int64_t NXg = p->NX_global > 0 ? p->NX_global : p->NX;
    int64_t NYg = p->NY_global > 0 ? p->NY_global : p->NY;
    acoustic_axis(p->NX + 2, p->DELTAX, p->origin_x, NXg, p->NPOINTS_PML, p->Rcoef, p->vp_ref,
                  p->USE_PML_XMIN, p->USE_PML_XMAX, sigma_x);
    acoustic_axis(p->NY + 2, p->DELTAY, p->origin_y, NYg, p->NPOINTS_PML, p->Rcoef, p->vp_ref,
                  p->USE_PML_YMIN, p->USE_PML_YMAX, tau_y);
    return ADS_OK;
}

int ads_acoustic_check_cfl(const ads_acoustic_params *p, double c_max)
{
    clear_error();
    if (!p) ADS_FAIL(ADS_ERR_ARG, "null params");
    double r = c_max * p->DELTAT *
               std::sqrt(1.0 / (p->DELTAX * p->DELTAX) + 1.0 / (p->DELTAY * p->DELTAY));
    if (!(r <= 1.0))
        ADS_FAIL(ADS_ERR_CFL, "CFL number %.4f > 1 (c_max=%g, dt=%g): the scheme is unstable", r,
                 c_max, p->DELTAT);
    return ADS_OK;
}

// compute_PML_Params(::ElasticPropagatorParams), src/Core.jl:231-407, one axis.
static void elastic_axis(int64_t len, double h, double origin, int64_t NXg, int64_t npml,
                         double npower, double kmax, double amax, double R, double vp_ref,
                         double dt, int use_min, int use_max, double *a, double *ah, double *b,
                         double *bh, double *K, double *Kh)
{
    const double thick = (double)npml * h;
    const double d0 = -(npower + 1) * vp_ref * std::log(R) / (2. * thick);
    const double oleft = thick, oright = ((double)NXg - 0.5) * h - thick;
    for (int64_t k = 0; k < len; k++) {
        double x = h * (double)k + origin;
        double d = 0, dh = 0, kk = 1, kh = 1, al = 0, alh = 0;
        if (use_min) {
            double ab = oleft - x;
            if (ab >= 0.0) {
                double an = ab / thick;
                d = d0 * std::pow(an, npower);
                kk = 1. + (kmax - 1.) * std::pow(an, npower);
                al = amax * an;
            }
            ab = oleft - (x + h / 2.);
            if (ab >= 0.0) {
                double an = ab / thick;
                dh = d0 * std::pow(an, npower);
                kh = 1. + (kmax - 1.) * std::pow(an, npower);
                alh = amax * an;
            }
        }
        if (use_max) {
            double ab = x - oright;
            if (ab >= 0) {
                double an = ab / thick;
                d = d0 * std::pow(an, npower);
                kk = 1. + (kmax - 1.) * std::pow(an, npower);
                al = amax * an;
            }
            ab = x + h / 2.0 - oright;
            if (ab >= 0) {
                double an = ab / thick;
                dh = d0 * std::pow(an, npower);
                kh = 1. + (kmax - 1.) * std::pow(an, npower);
                alh = amax * an;
            }
        }
        if (al < 0) al = 0;
        if (alh < 0) alh = 0;
        b[k] = std::exp(-(d / kk + al) * dt);
        bh[k] = std::exp(-(dh / kh + alh) * dt);
        a[k] = std::fabs(d) > 1e-6 ? d * (b[k] - 1.) / (kk * (d + kk * al)) : 0.0;
        ah[k] = std::fabs(dh) > 1e-6 ? dh * (bh[k] - 1.) / (kh * (dh + kh * alh)) : 0.0;
        K[k] = kk;
        Kh[k] = kh;
    }
}

int ads_elastic_pml_params(const ads_elastic_params *p, double *ax, double *bx, double *kx,
                           double *ay, double *by, double *ky)
{
    clear_error();
    if (!p || !ax || !bx || !kx || !ay || !by || !ky)
        ADS_FAIL(ADS_ERR_ARG, "ads_elastic_pml_params: null argument");
    if (p->NX < 3 || p->NY < 3) ADS_FAIL(ADS_ERR_SHAPE, "NX, NY must be >= 3");
    if (p->NPOINTS_PML < 1 || !(p->Rcoef > 0))
        ADS_FAIL(ADS_ERR_ARG, "bad PML parameters (NPOINTS_PML >= 1, Rcoef > 0)");
    int64_t NXg = p->NX_global > 0 ? p->NX_global : p->NX;
    int64_t NYg = p->NY_global > 0 ? p->NY_global : p->NY;
    elastic_axis(p->NX, p->DELTAX, p->origin_x, NXg, p->NPOINTS_PML, p->NPOWER, p->K_MAX_PML,
                 p->ALPHA_MAX_PML, p->Rcoef, p->vp_ref, p->DELTAT, p->USE_PML_XMIN,
                 p->USE_PML_XMAX, ax, ax + p->NX, bx, bx + p->NX, kx, kx + p->NX);
    elastic_axis(p->NY, p->DELTAY, p->origin_y, NYg, p->NPOINTS_PML, p->NPOWER, p->K_MAX_PML,
                 p->ALPHA_MAX_PML, p->Rcoef, p->vp_ref, p->DELTAT, p->USE_PML_YMIN,
                 p->USE_PML_YMAX, ay, ay + p->NY, by, by + p->NY, ky, ky + p->NY);
    return ADS_OK;
}

}  // extern "C"
