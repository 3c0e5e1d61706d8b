// ads_schedule.h -- checkpoint schedules for the adjoint sweeps and compact checkpoint records.
//
// A plan's time levels are cut into M segments of K levels.  The adjoint of segment m needs
// every level of it in the slot ring, rebuilt from the state at its start ("state m").  With S
// snapshot slots the reversal follows the binomial (Revolve-type) schedule of Griewank and
// Walther (ACM TOMS 26, 2000): to reverse n segments from a stored state with s free slots,
// advance j segments, store that state, reverse the n - j segments behind it with s - 1 slots,
// free the slot and reverse the first j segments with s slots.  Counting one advance per
// segment stepped (the rebuild of a segment before its adjoint included), the least number of
// advances is
//     cost(n, s) = r n - C(s + 1 + r, s + 2) + n,   r = min{ r : C(s + 1 + r, s + 1) >= n },
// (the initial zero state counts as one more snapshot), and
//     j = min(n - 1, max(1, C(s - 1 + r, s + 1), n - C(s + r, s)))
// is an optimal first advance.  Both were checked against an exhaustive dynamic program over
// every split for n < 120 and s < 16.  With S >= M - 1 the schedule is the one-level
// (sqrt(N)) scheme: one snapshot per segment start and every segment rebuilt once.
#pragma once

#include <vector>

#include "ads_common.h"

namespace ads {

// C(a + r, a), saturating at 2^62
inline int64_t binom_ar(int64_t a, int64_t r)
{
    if (r < 0 || a < 0) return 0;
    const int64_t cap = int64_t(1) << 62;
    int64_t v = 1;
    const int64_t k = std::min(a, r);
    const int64_t N = a + r;
    for (int64_t i = 1; i <= k; i++) {
        // v = C(N - k + i, i) built incrementally; v * (N - k + i) / i stays integral
        const int64_t num = N - k + i;
        if (v > cap / num) return cap;
        v = v * num / i;
    }
    return v;
}

// repetitions needed to reverse n segments with s free slots (s + 1 snapshots with the start)
inline int64_t sched_reps(int64_t n, int64_t s)
{
    int64_t r = 0;
    while (binom_ar(s + 1, r) < n) r++;
    return r;
}

// least number of segment advances to reverse n segments with s free slots
inline int64_t sched_cost(int64_t n, int64_t s)
{
    if (n <= 0) return 0;
    if (n == 1) return 1;
    if (s <= 0) return n * (n + 1) / 2;
    const int64_t r = sched_reps(n, s);
    return r * n - binom_ar(s + 2, r - 1) + n;
}

// optimal first advance for n >= 2 segments and s >= 1 free slots
inline int64_t sched_split(int64_t n, int64_t s)
{
    const int64_t r = sched_reps(n, s);
    const int64_t lo = r >= 2 ? binom_ar(s + 1, r - 2) : 1;
    return std::min(n - 1, std::max({int64_t(1), lo, n - binom_ar(s, r)}));
}

// The reversal over M segments with S slots.  Ops provides
//   int restore(int64_t m, int64_t q)  state m into the working slots (q < 0: m == 0, zeros)
//   int forward(int64_t m0, int64_t m1) advance from state m0 through segments m0 .. m1-1
//   int save(int64_t m, int64_t q)      the current state (state m) into slot q
//   int adjoint(int64_t m)              adjoint sweep of segment m (its levels are in the ring)
// The forward pass that recorded the traces has already run the first descent: it advanced
// through every segment and saved the states first_path() names, and the ring still holds the
// last segment.  `first` marks calls on that path.
struct Reversal {
    int64_t M = 0, S = 0;
    std::vector<int64_t> free;

    Reversal(int64_t M_, int64_t S_) : M(M_), S(S_) { reset(); }
    void reset()
    {
        free.clear();
        for (int64_t q = S - 1; q >= 0; q--) free.push_back(q);  // slot 0 is popped first
    }
    bool one_level() const { return S >= M - 1; }

    // (state, slot) pairs the forward pass saves, in level order
    std::vector<std::pair<int64_t, int64_t>> first_path()
    {
        reset();
        std::vector<std::pair<int64_t, int64_t>> path;
        if (one_level()) {
            for (int64_t m = 1; m < M; m++) path.push_back({m, m - 1});
            return path;
        }
        int64_t lo = 0;
        while (M - lo >= 2 && !free.empty()) {
            const int64_t j = lo + sched_split(M - lo, (int64_t)free.size());
            path.push_back({j, free.back()});
            free.pop_back();
            lo = j;
        }
        reset();
        return path;
    }

    template <class Ops> int run(Ops &op)
    {
        reset();
        if (M <= 0) return ADS_OK;
        if (one_level()) {  // the recursion below reduces to this loop; kept flat (no depth M)
            ADS_TRY(op.adjoint(M - 1));
            for (int64_t m = M - 2; m >= 0; m--) {
                ADS_TRY(op.restore(m, m - 1));
                ADS_TRY(op.forward(m, m + 1));
                ADS_TRY(op.adjoint(m));
            }
            return ADS_OK;
        }
        return rev(op, 0, M, -1, true);
    }

    template <class Ops> int rev(Ops &op, int64_t lo, int64_t hi, int64_t q_lo, bool first)
    {
        const int64_t n = hi - lo;
        if (n == 1) {
            if (!first) {
                ADS_TRY(op.restore(lo, q_lo));
                ADS_TRY(op.forward(lo, hi));
            }
            return op.adjoint(lo);
        }
        if (free.empty()) {  // no slot left: rebuild each segment from state lo
            for (int64_t t = hi - 1; t >= lo; t--) {
                if (!(first && t == hi - 1)) {
                    ADS_TRY(op.restore(lo, q_lo));
                    ADS_TRY(op.forward(lo, t + 1));
                }
                ADS_TRY(op.adjoint(t));
            }
            return ADS_OK;
        }
        const int64_t j = lo + sched_split(n, (int64_t)free.size());
        const int64_t q = free.back();
        free.pop_back();
        if (!first) {
            ADS_TRY(op.restore(lo, q_lo));
            ADS_TRY(op.forward(lo, j));
            ADS_TRY(op.save(j, q));
        }
        ADS_TRY(rev(op, j, hi, q, first));
        free.push_back(q);
        return rev(op, lo, j, q_lo, false);
    }
};

// Cost of a schedule in work units: the weighted advances beyond the forward pass
// (w[m] = work of segment m).
struct CostOps {
    const std::vector<double> *w;
    double extra = 0.0;
    int restore(int64_t, int64_t) { return ADS_OK; }
    int save(int64_t, int64_t) { return ADS_OK; }
    int adjoint(int64_t) { return ADS_OK; }
    int forward(int64_t m0, int64_t m1)
    {
        for (int64_t m = m0; m < m1; m++) extra += (*w)[m];
        return ADS_OK;
    }
};

// The (K, S) with the least modelled time whose workspace fits `budget`, over L levels cut into
// M = ceil(L / K) segments.  ws_size(K, S) is the workspace for interval K and S snapshot slots
// (S < 0: one per segment start); seg_work(K, M) the work of each segment (the shots active at
// its levels).  Model: the forward and the adjoint sweep pass every level once (an adjoint
// level costs adj_cost forward levels) and the schedule adds its rebuilds.  Among the one-level
// layouts that fit, every interval rebuilds all segments but the last, so the model prefers
// the longest last segment; otherwise it takes the most slots that fit.
template <class WsFn, class WorkFn>
inline bool choose_schedule(size_t budget, int64_t L, double total_work, double adj_cost,
                            WsFn ws_size, WorkFn seg_work, int64_t &Kout, int64_t &Sout)
{
    double best = 0.0;
    int64_t bK = 0, bS = 0, lastK = -1;
    L = std::max<int64_t>(L, 1);
    for (int64_t Mc = 1; Mc <= std::min<int64_t>(L, 4096); Mc++) {
        const int64_t K = (L + Mc - 1) / Mc;
        if (K == lastK) continue;
        lastK = K;
        const int64_t M = (L + K - 1) / K;
        if (ws_size(K, 0) > budget) continue;
        int64_t S = std::max<int64_t>(M - 1, 0);
        if (ws_size(K, -1) > budget) {  // the most slots that fit, below the one-level count
            int64_t lo = 0, hi = std::max<int64_t>(M - 2, 0);
            while (lo < hi) {
                const int64_t mid = (lo + hi + 1) / 2;
                if (ws_size(K, mid) <= budget) lo = mid;
                else hi = mid - 1;
            }
            S = lo;
        }
        if (S == 0 && M > 64) continue;  // quadratic rebuilds: never the fastest here
        const std::vector<double> w = seg_work(K, M);
        CostOps c{&w};
        Reversal rv(M, S);
        rv.run(c);
        const double t = (1.0 + adj_cost) * total_work + c.extra;
        if (bK == 0 || t < best * (1.0 - 1e-12)) {
            best = t;
            bK = K;
            bS = S;
        }
    }
    if (bK == 0) return false;
    Kout = bK;
    Sout = bS;
    return true;
}

// ---- compact records of the PML auxiliary fields ----
// The auxiliary fields (acoustic phi/psi, elastic CPML memory variables) are zero on every
// cell of the "quiet box" [qi0, qi1) x [qj0, qj1), where no damping profile acts; a record
// stores only the H x W cells outside it: the full rows above, the left and right runs of the
// rows beside, and the full rows below.
struct FrameBox {
    int H = 0, W = 0, qi0 = 0, qi1 = 0, qj0 = 0, qj1 = 0;
    __host__ __device__ int64_t cells() const
    {
        return (int64_t)H * W - (int64_t)(qi1 - qi0) * (qj1 - qj0);
    }
    // (i, j) -> its index f among the cells outside the box, or -1 inside the box (the
    // inverse of cell())
    __host__ __device__ int64_t index(int i, int j) const
    {
        const int64_t top = (int64_t)qi0 * W;
        if (i < qi0) return (int64_t)i * W + j;
        const int wm = qj0 + (W - qj1);
        const int64_t mid = (int64_t)(qi1 - qi0) * wm;
        if (i >= qi1) return top + mid + (int64_t)(i - qi1) * W + j;
        if (j >= qj0 && j < qj1) return -1;
        return top + (int64_t)(i - qi0) * wm + (j < qj0 ? j : qj0 + (j - qj1));
    }
    // f-th cell outside the box -> (i, j); f in [0, cells())
    __host__ __device__ void cell(int64_t f, int &i, int &j) const
    {
        const int64_t top = (int64_t)qi0 * W;
        if (f < top) {
            i = (int)(f / W);
            j = (int)(f - (int64_t)i * W);
            return;
        }
        f -= top;
        const int wl = qj0, wm = qj0 + (W - qj1);
        const int64_t mid = (int64_t)(qi1 - qi0) * wm;
        if (f < mid) {
            const int r = (int)(f / wm), c = (int)(f - (int64_t)r * wm);
            i = qi0 + r;
            j = c < wl ? c : qj1 + (c - wl);
            return;
        }
        f -= mid;
        const int r = (int)(f / W);
        i = qi1 + r;
        j = (int)(f - (int64_t)r * W);
    }
};

// The largest box on which both 1-D profiles vanish (rows where px[i] == 0, columns where
// py[j] == 0; each zero set is one run for the profiles used here).  An empty box (no such
// row or column, or a zero set that is not one run) makes every cell part of the record.
inline FrameBox quiet_box(const std::vector<double> &px, const std::vector<double> &py)
{
    FrameBox b;
    b.H = (int)px.size();
    b.W = (int)py.size();
    auto run = [](const std::vector<double> &v, int &lo, int &hi) {
        lo = hi = 0;
        int n = (int)v.size(), a = 0;
        while (a < n && v[a] != 0.0) a++;
        int e = a;
        while (e < n && v[e] == 0.0) e++;
        for (int k = e; k < n; k++)
            if (v[k] == 0.0) return false;  // a second zero run
        lo = a;
        hi = e;
        return e > a;
    };
    int i0, i1, j0, j1;
    if (!run(px, i0, i1) || !run(py, j0, j1)) {
        b.qi0 = b.qi1 = b.H;  // every row in the top band
        b.qj0 = b.qj1 = 0;
        return b;
    }
    b.qi0 = i0;
    b.qi1 = i1;
    b.qj0 = j0;
    b.qj1 = j1;
    return b;
}

// Pack (UNPACK = false) the cells outside the box of `act` shots of a pitched field
// (shot stride fstride, row pitch P) into rec[s * ncell + f], or unpack them back.
template <typename T, bool UNPACK>
__global__ __launch_bounds__(256) void k_frame_copy(T *__restrict__ field, int64_t fstride,
                                                    T *__restrict__ rec, FrameBox b, int P)
{
    const int64_t n = b.cells();
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const int s = blockIdx.y;
    int i, j;
    b.cell(f, i, j);
    T *cell = field + (int64_t)s * fstride + (int64_t)i * P + j;
    T *slot = rec + (int64_t)s * n + f;
    if (UNPACK) *cell = *slot;
    else *slot = *cell;
}

}  // namespace ads
