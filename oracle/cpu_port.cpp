// CPU port of the query + smoothing step, used only as the CPU baseline that
// bench.py times beside the GPU result (test infrastructure: nothing in the
// product loads it). Unlike oracle.cpp, which restates the reference
// step by step in FP64 for clarity, this is the code a CPU user of the
// reference would write for the same work: OpenMP over vertices and faces,
// fp32 positions, int32 CSRs over global ids.
//
// What it restates:
//   - VV and VF as CSRs: the relations rxmesh_queries.cuh:502-858 produces
//     (VV = both endpoints of every edge, VF = the three corners of every
//     face), built here by a counting sort over the edge list and fv, each
//     row sorted ascending;
//   - uniform Laplacian smoothing, Jacobi, x <- x + lambda (mean(N(v)) - x),
//     the smoothing the reference's apps run over VV.
// Timing: one warm-up run, then the median of `reps` runs (steady_clock),
// with `threads` OpenMP threads.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <vector>

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class F>
double median_time(int reps, F&& f)
{
    f();  // warm-up
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        const double a = now_s();
        f();
        t.push_back(now_s() - a);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

// Exclusive scan of cnt[0, n) into off[0, n] with a two-pass blocked scan.
void scan_parallel(const std::vector<int32_t>& cnt, int32_t* off, int64_t n)
{
    const int       T = omp_get_max_threads();
    std::vector<int64_t> part(T + 1, 0);
#pragma omp parallel num_threads(T)
    {
        const int     t = omp_get_thread_num();
        const int64_t b = n * t / T, e = n * (t + 1) / T;
        int64_t       s = 0;
        for (int64_t i = b; i < e; ++i) s += cnt[i];
        part[t + 1] = s;
#pragma omp barrier
#pragma omp single
        for (int i = 0; i < T; ++i) part[i + 1] += part[i];
        int64_t run = part[t];
        for (int64_t i = b; i < e; ++i) {
            off[i] = static_cast<int32_t>(run);
            run += cnt[i];
        }
    }
    off[n] = static_cast<int32_t>(part[T]);
}

// Counting sort of n (key, value) entries into nkeys rows: histogram and
// slot cursors with atomic increments, then every row sorted ascending.
template <class KeyF, class ValF>
void csr_by_key(int64_t n, int64_t nkeys, KeyF key, ValF val, int32_t* rp, int32_t* col,
                std::vector<int32_t>& cnt)
{
    cnt.assign(nkeys, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
#pragma omp atomic
        cnt[key(i)]++;
    }
    scan_parallel(cnt, rp, nkeys);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nkeys; ++k) cnt[k] = rp[k];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int32_t at;
#pragma omp atomic capture
        at = cnt[key(i)]++;
        col[at] = val(i);
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t k = 0; k < nkeys; ++k) std::sort(col + rp[k], col + rp[k + 1]);
}

}  // namespace

extern "C" {

// VV (from the E x 2 edge list ev) and VF (from the F x 3 face list fv) as
// CSRs over global ids. Returns the median seconds of one VV + VF build.
double cpu_query_vv_vf(int64_t nV, int64_t nE, int64_t nF, const uint32_t* ev, const uint32_t* fv, int threads,
                       int reps, int32_t* vv_rp, int32_t* vv_col, int32_t* vf_rp, int32_t* vf_col)
{
    omp_set_num_threads(threads);
    std::vector<int32_t> cnt;
    return median_time(reps, [&] {
        csr_by_key(
            2 * nE, nV, [&](int64_t i) { return ev[i]; },
            [&](int64_t i) { return static_cast<int32_t>(ev[i ^ 1]); }, vv_rp, vv_col, cnt);
        csr_by_key(
            3 * nF, nV, [&](int64_t i) { return fv[i]; }, [&](int64_t i) { return static_cast<int32_t>(i / 3); },
            vf_rp, vf_col, cnt);
    });
}

// `iters` Jacobi iterations of uniform Laplacian smoothing over the VV CSR,
// positions as fp32 SoA [3][nV]; the result is in out. Returns the median
// seconds of the whole run of iterations.
double cpu_smooth(int64_t nV, const int32_t* rp, const int32_t* col, const float* x, float lambda, int iters,
                  int threads, int reps, float* out, float* tmp)
{
    omp_set_num_threads(threads);
    return median_time(reps, [&] {
        const float* src = x;
        for (int it = 0; it < iters; ++it) {
            float* dst = ((iters - it) & 1) ? out : tmp;
#pragma omp parallel for schedule(static)
            for (int64_t v = 0; v < nV; ++v) {
                const int32_t b = rp[v], e = rp[v + 1];
                for (int c = 0; c < 3; ++c) {
                    const float* s = src + c * nV;
                    float        a = 0.f;
                    for (int32_t k = b; k < e; ++k) a += s[col[k]];
                    const float xv = s[v];
                    dst[c * nV + v] = (e > b) ? xv + lambda * (a / static_cast<float>(e - b) - xv) : xv;
                }
            }
            src = dst;
        }
    });
}

int cpu_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
