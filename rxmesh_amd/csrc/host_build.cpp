// Host connectivity build: see host_build.h.
#include "host_build.h"

#include <algorithm>
#include <cmath>
#include <atomic>
#include <unordered_map>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <numeric>
#include <thread>
#include <chrono>
#include <cstdio>

namespace rxm {

namespace {
thread_local unsigned tl_limit = 0;  // 0: no cap on this thread
}

unsigned num_threads()
{
    static unsigned n = [] {
        unsigned t = std::thread::hardware_concurrency();
        if (const char* s = std::getenv("RXM_NUM_THREADS")) t = static_cast<unsigned>(std::atoi(s));
        if (t < 1) t = 1;
        if (t > 64) t = 64;
        return t;
    }();
    return tl_limit ? std::min(n, tl_limit) : n;
}

ThreadLimit::ThreadLimit(unsigned k) : prev(tl_limit)
{
    if (k) tl_limit = k;
}
ThreadLimit::~ThreadLimit() { tl_limit = prev; }

namespace {

// Dynamic-chunk parallel loop over [0, n). f(i) must be thread-safe.
template <class F>
void parallel_for(size_t n, F&& f, size_t chunk = 256)
{
    unsigned T = num_threads();
    if (T == 1 || n <= chunk) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t>      next{0};
    std::vector<std::thread> pool;
    const unsigned           lim = tl_limit;
    for (unsigned t = 0; t < T; ++t)
        pool.emplace_back([&, lim] {
            ThreadLimit cap(lim);
            size_t s;
            while ((s = next.fetch_add(chunk)) < n) {
                size_t e = std::min(n, s + chunk);
                for (size_t i = s; i < e; ++i) f(i);
            }
        });
    for (auto& th : pool) th.join();
}

inline uint32_t round_up(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

// exclusive prefix sum of counts (n) into off (n + 1)
void exclusive_scan(const std::vector<uint32_t>& cnt, std::vector<uint32_t>& off)
{
    off.resize(cnt.size() + 1);
    off[0] = 0;
    for (size_t i = 0; i < cnt.size(); ++i) off[i + 1] = off[i] + cnt[i];
}

}  // namespace

// ---------------------------------------------------------------------------
// a1. reference: include/rxmesh/rxmesh.cpp:518-681. Edge ids are assigned in
// order of first appearance over the half-edges (f, i) -> (fv[i], fv[i+1]);
// instead of a hash map, half-edges are bucketed by their smaller endpoint
// and each bucket's groups are resolved to their first half-edge.
int supporting_prepare(HostMesh& m, const uint32_t* fv, uint32_t nF, uint32_t nV_hint)
{
    if (fv == nullptr || nF == 0) return kErrInvalidArg;
    if (3ull * nF >= kInvalid32) return kErrInvalidArg;
    m.nF = nF;
    m.fv.assign(fv, fv + 3ull * nF);
    uint32_t vmax = 0;
    for (uint64_t i = 0; i < 3ull * nF; ++i) vmax = std::max(vmax, m.fv[i]);
    if (vmax == kInvalid32) return kErrInvalidArg;
    m.nV = std::max(nV_hint, vmax + 1);
    for (uint32_t f = 0; f < nF; ++f) {
        const uint32_t* t = &m.fv[3ull * f];
        if (t[0] == t[1] || t[1] == t[2] || t[0] == t[2]) return kErrInvalidArg;
    }
    return kOk;
}

namespace {

// f(begin, end) over contiguous chunks of [0, n) on the build's threads
template <class F>
void parallel_chunks(size_t n, F&& f)
{
    const size_t nc = std::max<size_t>(1, std::min<size_t>(n / 65536 + 1, size_t(num_threads()) * 4));
    parallel_for(nc, [&](size_t c) { f(n * c / nc, n * (c + 1) / nc); }, 1);
}

void atomic_max(std::atomic<uint32_t>& a, uint32_t x)
{
    uint32_t cur = a.load(std::memory_order_relaxed);
    while (cur < x && !a.compare_exchange_weak(cur, x, std::memory_order_relaxed)) {
    }
}

}  // namespace

int supporting_finish(HostMesh& m)
{
    const uint64_t nH = 3ull * m.nF;
    const uint32_t nE = m.nE;
    if (m.ev.size() != 2ull * nE || m.fe_g.size() != nH || m.ef_off.size() != nE + 1ull ||
        m.vf_off.size() != m.nV + 1ull || m.ef_val.size() != nH || m.vf_val.size() != nH)
        return kErrInvalidArg;
    // consistent orientation: of the half-edges of an edge, only its first
    // (whose direction EV keeps) starts at ev[2e]
    {
        std::unique_ptr<std::atomic<uint8_t>[]> fwd(new std::atomic<uint8_t>[nE ? nE : 1u]());
        std::atomic<bool>                       oriented{true};
        parallel_chunks(nH, [&](size_t a, size_t b) {
            bool ok = true;
            for (size_t h = a; h < b; ++h) {
                const uint32_t e = m.fe_g[h];
                if (m.fv[h] == m.ev[2ull * e] && fwd[e].fetch_add(1, std::memory_order_relaxed) > 0) ok = false;
            }
            if (!ok) oriented.store(false, std::memory_order_relaxed);
        });
        m.oriented = oriented.load();
    }
    // statistics (rxmesh.cpp:589-660)
    std::unique_ptr<std::atomic<uint32_t>[]> val(new std::atomic<uint32_t>[m.nV ? m.nV : 1u]());
    std::atomic<bool>                        closed{true}, manifold{true};
    std::atomic<uint32_t>                    mef{0}, mval{0}, mfa{0};
    parallel_chunks(nE, [&](size_t a, size_t b) {
        bool     cl = true, mf = true;
        uint32_t mx = 0;
        for (size_t e = a; e < b; ++e) {
            const uint32_t c = m.ef_off[e + 1] - m.ef_off[e];
            if (c == 1) cl = false;
            if (c > 2) mf = false;
            mx = std::max(mx, c);
            val[m.ev[2 * e]].fetch_add(1, std::memory_order_relaxed);
            val[m.ev[2 * e + 1]].fetch_add(1, std::memory_order_relaxed);
        }
        if (!cl) closed.store(false, std::memory_order_relaxed);
        if (!mf) manifold.store(false, std::memory_order_relaxed);
        atomic_max(mef, mx);
    });
    parallel_chunks(m.nV, [&](size_t a, size_t b) {
        uint32_t mx = 0;
        for (size_t v = a; v < b; ++v) mx = std::max(mx, val[v].load(std::memory_order_relaxed));
        atomic_max(mval, mx);
    });
    parallel_chunks(m.nF, [&](size_t a, size_t b) {
        uint32_t mx = 0;
        for (size_t f = a; f < b; ++f) {
            uint32_t s = 0;
            for (int i = 0; i < 3; ++i) {
                const uint32_t e = m.fe_g[3 * f + i];
                s += m.ef_off[e + 1] - m.ef_off[e] - 1;
            }
            mx = std::max(mx, s);
        }
        atomic_max(mfa, mx);
    });
    m.closed         = closed.load();
    m.manifold       = manifold.load();
    m.max_edge_faces = mef.load();
    m.max_valence    = mval.load();
    m.max_face_adj   = mfa.load();
    return kOk;
}

int build_supporting(HostMesh& m, const uint32_t* fv, uint32_t nF, uint32_t nV_hint)
{
    int rc = supporting_prepare(m, fv, nF, nV_hint);
    if (rc != kOk) return rc;
    const uint64_t nH = 3ull * nF;

    // vertex -> faces, face order
    std::vector<uint32_t> cnt(m.nV, 0);
    for (uint64_t h = 0; h < nH; ++h) cnt[m.fv[h]]++;
    exclusive_scan(cnt, m.vf_off);
    m.vf_val.resize(nH);
    {
        std::vector<uint32_t> cur(m.vf_off.begin(), m.vf_off.end() - 1);
        for (uint64_t h = 0; h < nH; ++h) m.vf_val[cur[m.fv[h]]++] = static_cast<uint32_t>(h / 3);
    }

    // half-edges bucketed by min endpoint, in half-edge order
    auto he_a = [&](uint64_t h) { return m.fv[h]; };
    auto he_b = [&](uint64_t h) { return m.fv[(h % 3 == 2) ? h - 2 : h + 1]; };
    std::fill(cnt.begin(), cnt.end(), 0);
    for (uint64_t h = 0; h < nH; ++h) cnt[std::min(he_a(h), he_b(h))]++;
    std::vector<uint32_t> boff;
    exclusive_scan(cnt, boff);
    std::vector<uint32_t> bval(nH);
    {
        std::vector<uint32_t> cur(boff.begin(), boff.end() - 1);
        for (uint64_t h = 0; h < nH; ++h) bval[cur[std::min(he_a(h), he_b(h))]++] = static_cast<uint32_t>(h);
    }
    std::vector<uint32_t> leader(nH);
    parallel_for(m.nV, [&](size_t v) {
        for (uint32_t i = boff[v]; i < boff[v + 1]; ++i) {
            uint32_t hi = bval[i];
            uint32_t ki = std::max(he_a(hi), he_b(hi));
            uint32_t l  = hi;
            for (uint32_t j = boff[v]; j < i; ++j) {
                uint32_t hj = bval[j];
                if (std::max(he_a(hj), he_b(hj)) == ki) {
                    l = hj;
                    break;
                }
            }
            leader[hi] = l;
        }
    }, 4096);
    // edge id = rank of the leader half-edge
    std::vector<uint32_t> eid(nH, kInvalid32);
    uint32_t              nE = 0;
    for (uint64_t h = 0; h < nH; ++h)
        if (leader[h] == h) eid[h] = nE++;
    m.nE = nE;
    m.fe_g.resize(nH);
    m.ev.resize(2ull * nE);
    for (uint64_t h = 0; h < nH; ++h) {
        uint32_t e = eid[leader[h]];
        m.fe_g[h]  = e;
        if (leader[h] == h) {
            m.ev[2ull * e]     = he_a(h);
            m.ev[2ull * e + 1] = he_b(h);
        }
    }
    // edge -> faces, face order
    std::vector<uint32_t> ecnt(nE, 0);
    for (uint64_t h = 0; h < nH; ++h) ecnt[m.fe_g[h]]++;
    exclusive_scan(ecnt, m.ef_off);
    m.ef_val.resize(nH);
    {
        std::vector<uint32_t> cur(m.ef_off.begin(), m.ef_off.end() - 1);
        for (uint64_t h = 0; h < nH; ++h) m.ef_val[cur[m.fe_g[h]]++] = static_cast<uint32_t>(h / 3);
    }
    return supporting_finish(m);
}

// ---------------------------------------------------------------------------
// a2. Deterministic patchers. The reference grows patches with a GPU Lloyd
// iteration (patcher.cu:828-987) seeded from std::random_device, so its
// layout cannot be reproduced. The default is a recursive proportional
// bisection (Bisector below). The alternative keeps the reference's idea --
// k-means on the face graph: multi-source BFS assignment, re-centre each seed
// on its patch's most interior face, repeat -- with seeds taken at even
// strides of a BFS order, followed by halving of any patch above
// patch_size. Both use FIFO BFS with id tie-breaks: deterministic.
namespace {

struct FaceGraph
{
    std::vector<uint32_t> off, val;
};

// ---- level-synchronous BFS with the serial FIFO order ------------------------
// Threads of a BFS team meet at a barrier three times per level; a waiting
// thread spins briefly, then yields its core.
class SpinBarrier
{
   public:
    explicit SpinBarrier(unsigned n) : n_(n) {}
    void wait()
    {
        const unsigned g = gen_.load(std::memory_order_acquire);
        if (count_.fetch_add(1, std::memory_order_acq_rel) + 1 == n_) {
            count_.store(0, std::memory_order_relaxed);
            gen_.fetch_add(1, std::memory_order_release);
            return;
        }
        for (unsigned spins = 0; gen_.load(std::memory_order_acquire) == g; ++spins)
            if (spins > 256) std::this_thread::yield();
    }

   private:
    const unsigned        n_;
    std::atomic<unsigned> count_{0}, gen_{0};
};

// Threads a BFS over n faces may use: large sets only (a level costs three
// barriers), at most 16, shared with the `busy` other tasks in flight.
constexpr size_t kTeamMin = size_t(1) << 20;
unsigned bfs_team(size_t n, unsigned busy = 0)
{
    if (n < kTeamMin) return 1;
    return std::max(1u, std::min(16u, num_threads() / (busy + 1)));
}

// BFS from src over the faces f with ok(f) and dist[f] == kInvalid32: appends
// them to q in the order of the serial FIFO BFS and sets dist (0 at src).
// With nthr >= 2 each level runs on a team of threads with that same order:
// the serial BFS pops the level's faces in q order and pushes each one's
// unvisited neighbours in adjacency order, so a face reached at level L + 1
// belongs to the earliest level-L face in q that has it as a neighbour
// (claimed by an atomic max of (epoch, ~position): stale claims of older
// epochs lose, and no reset is needed), and the next level is the parents'
// claimed neighbours in (parent position, adjacency) order, placed by a
// prefix sum of per-thread counts. `cap` bounds the faces reached.
template <class Ok>
void fifo_bfs(const FaceGraph& g, uint32_t src, Ok ok, uint32_t* dist, std::atomic<uint64_t>* claim, uint32_t epoch,
              std::vector<uint32_t>& q, size_t cap, unsigned nthr)
{
    const size_t base = q.size();
    dist[src]         = 0;
    if (nthr < 2) {
        q.push_back(src);
        for (size_t h = base; h < q.size(); ++h) {
            const uint32_t f = q[h];
            for (uint32_t j = g.off[f]; j < g.off[f + 1]; ++j) {
                const uint32_t o = g.val[j];
                if (ok(o) && dist[o] == kInvalid32) {
                    dist[o] = dist[f] + 1;
                    q.push_back(o);
                }
            }
        }
        return;
    }
    q.resize(base + cap);
    q[base] = src;
    std::vector<size_t> cnt(nthr, 0);
    size_t              last = base + 1;
    SpinBarrier         bar(nthr);
    auto key = [&](size_t i) { return (uint64_t(epoch) << 32) | (0xFFFFFFFFu - static_cast<uint32_t>(i)); };
    // o at slot j of f's adjacency is f's to emit: claimed by f, first slot naming o
    auto mine = [&](uint32_t f, uint32_t j, uint32_t o, uint64_t k) {
        if (claim[o].load(std::memory_order_relaxed) != k) return false;
        for (uint32_t i = g.off[f]; i < j; ++i)
            if (g.val[i] == o) return false;
        return true;
    };
    auto body = [&](unsigned t) {
        size_t   lo = base, hi = base + 1;
        uint32_t lev = 0;
        while (lo < hi) {
            const size_t a = lo + (hi - lo) * t / nthr, b = lo + (hi - lo) * (t + 1) / nthr;
            for (size_t i = a; i < b; ++i) {
                const uint32_t f = q[i];
                const uint64_t k = key(i);
                for (uint32_t j = g.off[f]; j < g.off[f + 1]; ++j) {
                    const uint32_t o = g.val[j];
                    if (!ok(o) || dist[o] != kInvalid32) continue;
                    uint64_t cur = claim[o].load(std::memory_order_relaxed);
                    while (cur < k && !claim[o].compare_exchange_weak(cur, k, std::memory_order_relaxed)) {
                    }
                }
            }
            bar.wait();
            size_t c = 0;
            for (size_t i = a; i < b; ++i) {
                const uint32_t f = q[i];
                const uint64_t k = key(i);
                for (uint32_t j = g.off[f]; j < g.off[f + 1]; ++j) c += mine(f, j, g.val[j], k);
            }
            cnt[t] = c;
            bar.wait();
            size_t w = hi, tot = 0;
            for (unsigned u = 0; u < nthr; ++u) {
                if (u < t) w += cnt[u];
                tot += cnt[u];
            }
            for (size_t i = a; i < b; ++i) {
                const uint32_t f = q[i];
                const uint64_t k = key(i);
                for (uint32_t j = g.off[f]; j < g.off[f + 1]; ++j) {
                    const uint32_t o = g.val[j];
                    if (mine(f, j, o, k)) {
                        q[w++]  = o;
                        dist[o] = lev + 1;
                    }
                }
            }
            bar.wait();
            lo = hi;
            hi += tot;
            ++lev;
        }
        if (t == 0) last = hi;
    };
    std::vector<std::thread> team;
    const unsigned           lim = tl_limit;
    for (unsigned t = 1; t < nthr; ++t)
        team.emplace_back([&, t, lim] {
            ThreadLimit inherit(lim);
            body(t);
        });
    body(0);
    for (auto& th : team) th.join();
    q.resize(last);
}

FaceGraph face_graph(const HostMesh& m)
{
    FaceGraph g;
    g.off.assign(m.nF + 1, 0);
    for (uint32_t f = 0; f < m.nF; ++f) {
        uint32_t c = 0;
        for (int i = 0; i < 3; ++i) {
            uint32_t e = m.fe_g[3ull * f + i];
            c += m.ef_off[e + 1] - m.ef_off[e] - 1;
        }
        g.off[f + 1] = g.off[f] + c;
    }
    g.val.resize(g.off[m.nF]);
    parallel_for(m.nF, [&](size_t f) {
        uint32_t k = g.off[f];
        for (int i = 0; i < 3; ++i) {
            uint32_t e = m.fe_g[3ull * f + i];
            for (uint32_t j = m.ef_off[e]; j < m.ef_off[e + 1]; ++j)
                if (m.ef_val[j] != f) g.val[k++] = m.ef_val[j];
        }
    }, 8192);
    return g;
}

// Multi-source BFS from seeds; faces not reached get fresh seeds (smallest
// unassigned face id). Returns the number of patches.
uint32_t bfs_assign(const FaceGraph& g, uint32_t nF, std::vector<uint32_t>& seeds,
                    std::vector<uint32_t>& fp)
{
    fp.assign(nF, kInvalid32);
    std::vector<uint32_t> q;
    q.reserve(nF);
    for (uint32_t k = 0; k < seeds.size(); ++k) {
        fp[seeds[k]] = k;
        q.push_back(seeds[k]);
    }
    size_t   head = 0;
    uint32_t scan = 0;
    while (true) {
        while (head < q.size()) {
            uint32_t f = q[head++];
            for (uint32_t j = g.off[f]; j < g.off[f + 1]; ++j) {
                uint32_t n = g.val[j];
                if (fp[n] == kInvalid32) {
                    fp[n] = fp[f];
                    q.push_back(n);
                }
            }
        }
        while (scan < nF && fp[scan] != kInvalid32) ++scan;
        if (scan == nF) break;
        fp[scan] = static_cast<uint32_t>(seeds.size());
        seeds.push_back(scan);
        q.push_back(scan);
    }
    return static_cast<uint32_t>(seeds.size());
}

// New seed of each patch = its face farthest from the patch border.
void recenter(const FaceGraph& g, uint32_t nF, const std::vector<uint32_t>& fp,
              std::vector<uint32_t>& seeds)
{
    std::vector<uint32_t> dist(nF, kInvalid32), q;
    q.reserve(nF);
    for (uint32_t f = 0; f < nF; ++f) {
        bool border = (g.off[f + 1] - g.off[f]) < 3;
        for (uint32_t j = g.off[f]; j < g.off[f + 1] && !border; ++j)
            border = fp[g.val[j]] != fp[f];
        if (border) {
            dist[f] = 0;
            q.push_back(f);
        }
    }
    size_t head = 0;
    while (head < q.size()) {
        uint32_t f = q[head++];
        for (uint32_t j = g.off[f]; j < g.off[f + 1]; ++j) {
            uint32_t n = g.val[j];
            if (fp[n] == fp[f] && dist[n] == kInvalid32) {
                dist[n] = dist[f] + 1;
                q.push_back(n);
            }
        }
    }
    std::vector<uint32_t> best(seeds.size(), 0);
    for (uint32_t f = 0; f < nF; ++f) {
        uint32_t k = fp[f];
        uint32_t d = (dist[f] == kInvalid32) ? 0 : dist[f];
        if (d > best[k] || (d == best[k] && dist[seeds[k]] != kInvalid32 && d > 0 && f < seeds[k])) {
            best[k]  = d;
            seeds[k] = f;
        }
    }
}

// Recursive proportional bisection: a face set that must become k patches is
// cut in two at the median of (d_a - d_b), where a and b are two mutually far
// faces of the set (BFS distances inside the set), into parts of
// round(n * k1 / k) and the rest faces for k1 = k / 2 and k - k1 patches.
// With k = ceil(F / cap) at the root every leaf holds about F / k <= cap
// faces, the fewest patches the cap allows. Leaves are listed depth-first
// (neighbouring patches get nearby ids); ties break on distances and face
// ids, so the result is deterministic. Large independent subtrees run on
// worker threads: each set in flight carries its own tag, and a task reads
// or writes the per-face scratch of its own faces only.
class Bisector
{
   public:
    using Leaves = std::vector<std::vector<uint32_t>>;

    // orig: the caller's id of each face of g (null: g is numbered as the
    // caller's faces); every tie breaks on these ids, so a relabelled graph
    // gives the patches of the original numbering
    Bisector(const FaceGraph& g, uint32_t nF, const std::vector<uint32_t>* orig = nullptr)
        : g_(g), orig_(orig), tag_(new std::atomic<uint32_t>[nF]), dist_(nF, kInvalid32), side_(nF, 0)
    {
        for (auto& cs : cside_) cs.assign(nF, 0);
        for (uint32_t f = 0; f < nF; ++f) tag_[f].store(kInvalid32, std::memory_order_relaxed);
    }

    // Entry: every connected component gets ceil(size / cap) patches; fp
    // receives patch ids in leaf order, the return value is the patch count.
    uint32_t run_all(uint32_t nF, uint32_t cap, std::vector<uint32_t>& fp)
    {
        std::vector<uint32_t> all(nF);
        std::iota(all.begin(), all.end(), 0u);
        uint32_t next = 0;
        for (auto& c : components(all)) {
            std::sort(c.begin(), c.end());
            const uint32_t k = static_cast<uint32_t>((c.size() + cap - 1) / cap);
            // on a graph relabelled in BFS order a component is a run of ids
            // that the relabelling BFS swept from its smallest caller id, so
            // the root split's first BFS would end at the run's last id
            const uint32_t far = orig_ ? c.back() : kInvalid32;
            for (const auto& leaf : run(std::move(c), k, cap, far)) {
                for (uint32_t f : leaf) fp[f] = next;
                ++next;
            }
        }
        return next;
    }

   private:
    static constexpr size_t kParallelMin = size_t(1) << 15;

    // far: the face a BFS from the set's smallest caller id reaches last, when
    // already known (kInvalid32 otherwise)
    Leaves run(std::vector<uint32_t> faces, uint32_t k, uint32_t cap, uint32_t far = kInvalid32)
    {
        if (k <= 1 && faces.size() <= cap) return Leaves{std::move(faces)};
        if (k <= 1) k = 2;
        std::vector<uint32_t> A, B;
        uint32_t              kA, kB;
        split(std::move(faces), k, cap, A, B, kA, kB, far);
        Leaves left, right;
        if (A.size() + B.size() >= kParallelMin && workers_.fetch_add(1) + 1 < num_threads()) {
            const unsigned lim = tl_limit;
            std::thread th([&, lim] {
                ThreadLimit inherit(lim);
                left = run(std::move(A), kA, cap);
            });
            right = run(std::move(B), kB, cap);
            th.join();
            workers_.fetch_sub(1);
        } else {
            if (A.size() + B.size() >= kParallelMin) workers_.fetch_sub(1);
            left  = run(std::move(A), kA, cap);
            right = run(std::move(B), kB, cap);
        }
        for (auto& l : right) left.push_back(std::move(l));
        return left;
    }

    void split(std::vector<uint32_t> faces, uint32_t k, uint32_t cap, std::vector<uint32_t>& A,
               std::vector<uint32_t>& B, uint32_t& kA, uint32_t& kB, uint32_t far = kInvalid32)
    {
        const uint32_t n  = static_cast<uint32_t>(faces.size());
        const uint32_t k1 = k / 2, k2 = k - k1;
        const uint32_t n1 = static_cast<uint32_t>((uint64_t(n) * k1 + k / 2) / k);
        const uint32_t tag = mark(faces);
        // three candidate cut axes between mutually far faces a, b, c; the
        // one whose split cuts the fewest face adjacencies wins (keeps cuts
        // short on meshes with handles, e.g. around a torus tube)
        std::vector<uint32_t> q;
        uint32_t start = faces[0];  // the set's smallest caller id
        {
            const size_t          nc = chunks(n);
            std::vector<uint32_t> cs(nc, faces[0]);
            each_chunk(n, nc, [&](size_t c, size_t a, size_t b) {
                for (size_t i = a; i < b; ++i)
                    if (oid(faces[i]) < oid(cs[c])) cs[c] = faces[i];
            });
            for (uint32_t f : cs)
                if (oid(f) < oid(start)) start = f;
        }
        const uint32_t a = (far != kInvalid32) ? far : farthest(faces, start, tag, q);
        std::vector<uint32_t> d[3];
        const uint32_t        b = farthest(faces, a, tag, q);
        d[0]                    = dists(faces);
        bfs(faces, b, tag, q);
        d[1]           = dists(faces);
        // the face farthest from both (largest min distance, ties: smallest
        // caller id); per chunk, then the chunks in order by the same rule
        uint32_t c     = start;
        uint32_t cbest = 0;
        {
            const size_t          nc = chunks(n);
            std::vector<uint32_t> cc(nc, start), cm(nc, 0);
            each_chunk(n, nc, [&](size_t k, size_t a, size_t b) {
                for (size_t i = a; i < b; ++i) {
                    const uint32_t m = std::min(d[0][i], d[1][i]);
                    if (m != kInvalid32 && (m > cm[k] || (m == cm[k] && m > 0 && oid(faces[i]) < oid(cc[k])))) {
                        cm[k] = m;
                        cc[k] = faces[i];
                    }
                }
            });
            for (size_t k = 0; k < nc; ++k)
                if (cm[k] > cbest || (cm[k] == cbest && cm[k] > 0 && oid(cc[k]) < oid(c))) {
                    cbest = cm[k];
                    c     = cc[k];
                }
        }
        bfs(faces, c, tag, q);
        d[2] = dists(faces);
        // (u, w, z): cut along d_u - d_w; faces tied on it (a level set)
        // are ordered by the distance to the third face z, so the median
        // splits a tie band into two connected runs, then by face id
        const int pairs[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
        uint64_t  cuts[3];
        auto      evaluate = [&](int c) {
            const auto& pr = pairs[c];
            select_side(faces, n1, d[pr[0]], d[pr[1]], d[pr[2]], cside_[c]);
            uint64_t cut = 0;
            for (uint32_t f : faces) {
                if (cside_[c][f] != 1u) continue;
                for (uint32_t j = g_.off[f]; j < g_.off[f + 1]; ++j) {
                    const uint32_t o = g_.val[j];
                    cut += (in(o, tag) && cside_[c][o] == 2u);
                }
            }
            cuts[c] = cut;
        };
        if (n >= kParallelMin * 8 && num_threads() >= 3) {  // large sets: the three candidates on their own threads
            std::thread t1(evaluate, 1), t2(evaluate, 2);
            evaluate(0);
            t1.join();
            t2.join();
        } else {
            for (int c = 0; c < 3; ++c) evaluate(c);
        }
        int best = 0;
        for (int c = 1; c < 3; ++c)
            if (cuts[c] < cuts[best]) best = c;
        // sides of the best cut (1 = A, 2 = B); each half is rebuilt in
        // ascending face order by one pass over `faces` (sorted on entry)
        // instead of sorting it
        each(n, [&](size_t i) { side_[faces[i]] = cside_[best][faces[i]]; });
        auto collect = [&] {
            if (n < kTeamMin) {
                A.clear();
                B.clear();
                for (uint32_t f : faces) (side_[f] == 1u ? A : B).push_back(f);
                return;
            }
            // large sets: per-chunk counts, then each chunk writes its run of A
            // and of B (the same ascending order as the serial pass)
            const size_t          nc = std::min<size_t>(64, n / 65536 + 1);
            std::vector<uint32_t> ca(nc + 1, 0), cb(nc + 1, 0);
            parallel_for(nc, [&](size_t c) {
                uint32_t k = 0;
                for (size_t i = n * c / nc; i < n * (c + 1) / nc; ++i) k += side_[faces[i]] == 1u;
                ca[c + 1] = k;
                cb[c + 1] = static_cast<uint32_t>(n * (c + 1) / nc - n * c / nc) - k;
            }, 1);
            for (size_t c = 0; c < nc; ++c) {
                ca[c + 1] += ca[c];
                cb[c + 1] += cb[c];
            }
            A.resize(ca[nc]);
            B.resize(cb[nc]);
            parallel_for(nc, [&](size_t c) {
                uint32_t ia = ca[c], ib = cb[c];
                for (size_t i = n * c / nc; i < n * (c + 1) / nc; ++i) {
                    const uint32_t f = faces[i];
                    if (side_[f] == 1u) A[ia++] = f;
                    else B[ib++] = f;
                }
            }, 1);
        };
        collect();
        // keep both halves connected: pieces cut off from the larger
        // component of one half border the other half, so they move there
        if (n >= kTeamMin && num_threads() >= 2) {
            // both halves' components at once (disjoint sets: own tags and
            // scratch); B's again only where A gave pieces away
            Leaves         ca, cb;
            const unsigned lim = tl_limit;
            std::thread    th([&, lim] {
                ThreadLimit inherit(lim);
                cb = components(B);
            });
            ca = components(A);
            th.join();
            for (size_t i = 1; i < ca.size(); ++i)
                for (uint32_t f : ca[i]) side_[f] = 2u;
            if (ca.size() > 1) {
                collect();
                cb = components(B);
            }
            for (size_t i = 1; i < cb.size(); ++i)
                for (uint32_t f : cb[i]) side_[f] = 1u;
            collect();
        } else {
            move_cut_off(A, 2u);
            collect();
            move_cut_off(B, 1u);
            collect();
        }
        each(n, [&](size_t i) { side_[faces[i]] = 0; });
        std::vector<uint32_t>().swap(faces);
        kA = std::max<uint32_t>(k1, static_cast<uint32_t>((A.size() + cap - 1) / cap));
        kB = std::max<uint32_t>(k2, static_cast<uint32_t>((B.size() + cap - 1) / cap));
    }

    bool in(uint32_t f, uint32_t tag) const { return tag_[f].load(std::memory_order_relaxed) == tag; }
    uint32_t oid(uint32_t f) const { return orig_ ? (*orig_)[f] : f; }

    // Side 1 for the n1 faces first in the order (du - dw, dz, face id), side
    // 2 for the rest, written to side[f]. The key is an integer of small
    // range, so a histogram finds the key value the median falls on and only
    // the faces tied on it are ordered (by dz, then face id): the same
    // partition a full sort gives. Faces not reached from u (another
    // component) sort last; faces not reached from w count distance 0.
    void select_side(const std::vector<uint32_t>& faces, uint32_t n1, const std::vector<uint32_t>& du,
                                const std::vector<uint32_t>& dw, const std::vector<uint32_t>& dz,
                                std::vector<uint8_t>& side)
    {
        const uint32_t n = static_cast<uint32_t>(faces.size());
        int64_t        lo = INT64_MAX, hi = INT64_MIN;
        auto           key = [&](uint32_t i) {
            const int64_t x = (du[i] == kInvalid32) ? (int64_t(1) << 40) : int64_t(du[i]);
            const int64_t y = (dw[i] == kInvalid32) ? 0 : int64_t(dw[i]);
            return x - y;
        };
        bool far_keys = false;
        for (uint32_t i = 0; i < n; ++i) {
            if (du[i] == kInvalid32) {
                far_keys = true;
                continue;
            }
            const int64_t k = key(i);
            lo              = std::min(lo, k);
            hi              = std::max(hi, k);
        }
        if (far_keys || n == 0) {  // general path: one selection over all keys
            std::vector<std::pair<std::pair<int64_t, uint64_t>, uint32_t>> kv(n);
            for (uint32_t i = 0; i < n; ++i) kv[i] = {{key(i), (uint64_t(dz[i]) << 32) | oid(faces[i])}, faces[i]};
            std::nth_element(kv.begin(), kv.begin() + n1, kv.end());
            for (uint32_t i = 0; i < n; ++i) side[kv[i].second] = (i < n1) ? 1u : 2u;
            return;
        }
        std::vector<uint32_t> hist(static_cast<size_t>(hi - lo + 1), 0);
        for (uint32_t i = 0; i < n; ++i) hist[static_cast<size_t>(key(i) - lo)]++;
        size_t   kb    = 0;
        uint32_t below = 0;
        while (below + hist[kb] <= n1 && kb + 1 < hist.size()) below += hist[kb++];
        const int64_t         K = lo + static_cast<int64_t>(kb);
        std::vector<std::pair<uint64_t, uint32_t>> tie;
        for (uint32_t i = 0; i < n; ++i) {
            const int64_t k = key(i);
            if (k < K) side[faces[i]] = 1u;
            else if (k > K) side[faces[i]] = 2u;
            else tie.push_back({(uint64_t(dz[i]) << 32) | oid(faces[i]), faces[i]});
        }
        const uint32_t take = n1 - below;  // tied faces that go to side 1
        if (take > 0 && take < tie.size()) std::nth_element(tie.begin(), tie.begin() + take, tie.end());
        for (size_t t = 0; t < tie.size(); ++t) side[tie[t].second] = (t < take) ? 1u : 2u;
    }

    // loops over a face set: on all threads for sets above 2^20 faces
    static size_t chunks(size_t n) { return n >= kTeamMin ? std::min<size_t>(64, n / 65536 + 1) : 1; }
    template <class F>
    static void each_chunk(size_t n, size_t nc, F&& f)
    {
        if (nc == 1) {
            f(size_t(0), size_t(0), n);
            return;
        }
        parallel_for(nc, [&](size_t c) { f(c, n * c / nc, n * (c + 1) / nc); }, 1);
    }
    template <class F>
    static void each(size_t n, F&& f)
    {
        each_chunk(n, chunks(n), [&](size_t, size_t a, size_t b) {
            for (size_t i = a; i < b; ++i) f(i);
        });
    }

    uint32_t mark(const std::vector<uint32_t>& faces)
    {
        const uint32_t tag = next_tag_.fetch_add(1);
        each(faces.size(), [&](size_t i) { tag_[faces[i]].store(tag, std::memory_order_relaxed); });
        return tag;
    }

    // connected components of a face set, largest first (ties: smaller first face)
    Leaves components(const std::vector<uint32_t>& faces)
    {
        const uint32_t tag = mark(faces);
        for (uint32_t f : faces) dist_[f] = kInvalid32;
        Leaves out;
        for (uint32_t s : faces) {
            if (dist_[s] != kInvalid32) continue;
            std::vector<uint32_t> c{s};
            dist_[s] = 0;
            for (size_t h = 0; h < c.size(); ++h) {
                const uint32_t f = c[h];
                for (uint32_t j = g_.off[f]; j < g_.off[f + 1]; ++j) {
                    const uint32_t o = g_.val[j];
                    if (in(o, tag) && dist_[o] == kInvalid32) {
                        dist_[o] = 0;
                        c.push_back(o);
                    }
                }
            }
            out.push_back(std::move(c));
        }
        // largest first; equal sizes in the order of their smallest caller id
        std::vector<uint32_t> first(out.size());
        for (size_t k = 0; k < out.size(); ++k) {
            uint32_t mn = kInvalid32;
            for (uint32_t f : out[k]) mn = std::min(mn, oid(f));
            first[k] = mn;
        }
        std::vector<size_t> ord(out.size());
        std::iota(ord.begin(), ord.end(), size_t(0));
        std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
            return out[x].size() != out[y].size() ? out[x].size() > out[y].size() : first[x] < first[y];
        });
        Leaves sorted;
        sorted.reserve(out.size());
        for (size_t k : ord) sorted.push_back(std::move(out[k]));
        return sorted;
    }

    // every component of the (sorted) set but the largest changes side
    void move_cut_off(const std::vector<uint32_t>& set, uint32_t to_side)
    {
        auto comps = components(set);
        for (size_t i = 1; i < comps.size(); ++i)
            for (uint32_t f : comps[i]) side_[f] = to_side;
    }

    // BFS inside the set tagged `tag`; fills dist_ for its faces
    void bfs(const std::vector<uint32_t>& faces, uint32_t src, uint32_t tag, std::vector<uint32_t>& q)
    {
        for (uint32_t f : faces) dist_[f] = kInvalid32;
        q.clear();
        q.push_back(src);
        dist_[src] = 0;
        for (size_t h = 0; h < q.size(); ++h) {
            const uint32_t f = q[h];
            for (uint32_t j = g_.off[f]; j < g_.off[f + 1]; ++j) {
                const uint32_t o = g_.val[j];
                if (in(o, tag) && dist_[o] == kInvalid32) {
                    dist_[o] = dist_[f] + 1;
                    q.push_back(o);
                }
            }
        }
    }
    // last face reached by a BFS from src (largest distance; FIFO order breaks ties)
    uint32_t farthest(const std::vector<uint32_t>& faces, uint32_t src, uint32_t tag, std::vector<uint32_t>& q)
    {
        bfs(faces, src, tag, q);
        return q.back();
    }

    std::vector<uint32_t> dists(const std::vector<uint32_t>& faces) const
    {
        std::vector<uint32_t> r(faces.size());
        each(faces.size(), [&](size_t i) { r[i] = dist_[faces[i]]; });
        return r;
    }

    const FaceGraph&                         g_;
    const std::vector<uint32_t>*             orig_;
    std::unique_ptr<std::atomic<uint32_t>[]> tag_;
    std::vector<uint32_t>                    dist_, side_;
    std::vector<uint8_t>                     cside_[3];  // sides of the three candidate cuts
    std::atomic<uint32_t>                    next_tag_{0};
    std::atomic<uint32_t>                    workers_{0};
};

// BFS distances restricted to a face subset (mask == tag)
void subset_bfs(const FaceGraph& g, const std::vector<uint32_t>& faces, uint32_t src,
                const std::vector<uint32_t>& tag_of, uint32_t tag,
                std::vector<uint32_t>& dist_out)
{
    // dist_out is indexed by position in `faces`; use a local map
    std::vector<std::pair<uint32_t, uint32_t>> pos;
    pos.reserve(faces.size());
    for (uint32_t i = 0; i < faces.size(); ++i) pos.push_back({faces[i], i});
    std::sort(pos.begin(), pos.end());
    auto index_of = [&](uint32_t f) {
        auto it = std::lower_bound(pos.begin(), pos.end(), std::make_pair(f, 0u));
        return it->second;
    };
    dist_out.assign(faces.size(), kInvalid32);
    std::vector<uint32_t> q{src};
    dist_out[index_of(src)] = 0;
    size_t head = 0;
    while (head < q.size()) {
        uint32_t f  = q[head++];
        uint32_t df = dist_out[index_of(f)];
        for (uint32_t j = g.off[f]; j < g.off[f + 1]; ++j) {
            uint32_t n = g.val[j];
            if (tag_of[n] != tag) continue;
            uint32_t in = index_of(n);
            if (dist_out[in] == kInvalid32) {
                dist_out[in] = df + 1;
                q.push_back(n);
            }
        }
    }
}

}  // namespace

int run_patcher(HostMesh& m, uint32_t patch_size, int mode)
{
    if (patch_size < 16) return kErrInvalidArg;
    m.patch_size = patch_size;
    const uint32_t nF = m.nF;
    std::vector<uint32_t> fp;
    if (nF <= patch_size) {
        fp.assign(nF, 0);
        return set_patches(m, fp.data(), 1, nullptr, nullptr);
    }
    auto clk = [] { return std::chrono::steady_clock::now(); };
    auto t0  = clk();
    FaceGraph g = face_graph(m);
    auto t1 = clk();
    // BFS order over all components from face 0 (on a team of threads for a
    // large mesh, in the same order)
    std::vector<uint32_t> order, rank(nF, kInvalid32);
    {
        std::vector<uint32_t>                    lvl(nF, kInvalid32);
        std::unique_ptr<std::atomic<uint64_t>[]> claim(new std::atomic<uint64_t>[nF]());
        order.reserve(nF);
        for (uint32_t s = 0; s < nF; ++s) {
            if (lvl[s] != kInvalid32) continue;
            const size_t left = nF - order.size();
            fifo_bfs(g, s, [](uint32_t) { return true; }, lvl.data(), claim.get(), 1u, order, left, bfs_team(left));
        }
        parallel_for(nF, [&](size_t i) { rank[order[i]] = static_cast<uint32_t>(i); }, size_t(1) << 16);
    }
    uint32_t nP = 0;
    if (mode == kPatcherKMeans) {
        const uint32_t P0 = (nF + patch_size - 1) / patch_size;
        std::vector<uint32_t> seeds(P0);
        for (uint32_t k = 0; k < P0; ++k) seeds[k] = order[(uint64_t(k) * nF) / P0];
        const int iters = 4;
        for (int it = 0; it < iters; ++it) {
            bfs_assign(g, nF, seeds, fp);
            if (it + 1 < iters) recenter(g, nF, fp, seeds);
        }
        nP = static_cast<uint32_t>(seeds.size());
    } else {
        // bisection over the faces relabelled in BFS order: neighbouring
        // faces get nearby ids, so its many BFS passes stay in cache however
        // the input is numbered; the patches map back to the input ids
        FaceGraph gr;
        gr.off.assign(nF + 1, 0);
        for (uint32_t i = 0; i < nF; ++i) gr.off[i + 1] = gr.off[i] + (g.off[order[i] + 1] - g.off[order[i]]);
        gr.val.resize(g.val.size());
        parallel_for(nF, [&](size_t i) {
            const uint32_t f = order[i];
            uint32_t       k = gr.off[i];
            for (uint32_t j = g.off[f]; j < g.off[f + 1]; ++j) gr.val[k++] = rank[g.val[j]];
        }, 8192);
        auto t2 = clk();
        std::vector<uint32_t> fpr(nF, kInvalid32);
        Bisector bs(gr, nF, &order);
        nP = bs.run_all(nF, patch_size, fpr);
        auto t3 = clk();
        fp.assign(nF, kInvalid32);
        for (uint32_t i = 0; i < nF; ++i) fp[order[i]] = fpr[i];
        if (std::getenv("RXM_TIMING")) {
            auto s_ = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
            std::fprintf(stderr, "patcher: face graph %.3f s, relabel %.3f s, bisection %.3f s\n", s_(t0, t1),
                         s_(t1, t2), s_(t2, t3));
        }
    }
    // recursive bisection of oversize patches
    std::vector<std::vector<uint32_t>> members(nP);
    for (uint32_t f = 0; f < nF; ++f) members[fp[f]].push_back(f);
    for (size_t k = 0; k < members.size(); ++k) {
        while (members[k].size() > patch_size) {
            auto& faces = members[k];
            std::vector<uint32_t> d;
            subset_bfs(g, faces, faces[0], fp, static_cast<uint32_t>(k), d);
            // a = farthest from faces[0], b = farthest from a (ties: first)
            auto far = [&](const std::vector<uint32_t>& dd) {
                uint32_t bi = 0;
                for (uint32_t i = 0; i < dd.size(); ++i)
                    if (dd[i] != kInvalid32 && (dd[bi] == kInvalid32 || dd[i] > dd[bi])) bi = i;
                return bi;
            };
            uint32_t a = faces[far(d)];
            std::vector<uint32_t> da, db;
            subset_bfs(g, faces, a, fp, static_cast<uint32_t>(k), da);
            uint32_t b = faces[far(da)];
            subset_bfs(g, faces, b, fp, static_cast<uint32_t>(k), db);
            std::vector<std::pair<int64_t, uint32_t>> key(faces.size());
            for (uint32_t i = 0; i < faces.size(); ++i) {
                int64_t x = (da[i] == kInvalid32) ? (1 << 30) : da[i];
                int64_t y = (db[i] == kInvalid32) ? (1 << 30) : db[i];
                key[i] = {x - y, faces[i]};
            }
            std::sort(key.begin(), key.end());
            size_t half = faces.size() / 2;
            std::vector<uint32_t> A, B;
            for (size_t i = 0; i < key.size(); ++i) (i < half ? A : B).push_back(key[i].second);
            std::sort(A.begin(), A.end());
            std::sort(B.begin(), B.end());
            uint32_t nk = static_cast<uint32_t>(members.size());
            for (uint32_t f : B) fp[f] = nk;
            members[k] = A;
            members.push_back(B);  // invalidates `faces`; loop re-reads members[k]
        }
    }
    // renumber patches by the BFS rank of their earliest face (locality of
    // the patch-linear attribute order)
    const uint32_t P = static_cast<uint32_t>(members.size());
    std::vector<std::pair<uint32_t, uint32_t>> first(P, {kInvalid32, 0});
    for (uint32_t k = 0; k < P; ++k) {
        first[k].second = k;
        for (uint32_t f : members[k]) first[k].first = std::min(first[k].first, rank[f]);
    }
    std::sort(first.begin(), first.end());
    std::vector<uint32_t> remap(P);
    for (uint32_t i = 0; i < P; ++i) remap[first[i].second] = i;
    for (uint32_t f = 0; f < nF; ++f) fp[f] = remap[fp[f]];
    return set_patches(m, fp.data(), P, nullptr, nullptr);
}

int set_patches(HostMesh& m, const uint32_t* face_patch, uint32_t P, const uint32_t* patches_val,
                const uint32_t* patches_offset)
{
    if (!face_patch || P == 0) return kErrInvalidArg;
    m.P = P;
    m.face_patch.assign(face_patch, face_patch + m.nF);
    std::vector<uint32_t> cnt(P, 0);
    for (uint32_t f = 0; f < m.nF; ++f) {
        if (m.face_patch[f] >= P) return kErrInvalidArg;
        cnt[m.face_patch[f]]++;
    }
    for (uint32_t p = 0; p < P; ++p)
        if (cnt[p] == 0) return kErrInvalidArg;  // empty patch
    if (patches_val && patches_offset) {
        m.patches_val.assign(patches_val, patches_val + m.nF);
        m.patches_offset.assign(patches_offset, patches_offset + P);
        uint32_t s = 0;
        for (uint32_t p = 0; p < P; ++p) {
            if (m.patches_offset[p] < s || m.patches_offset[p] > m.nF) return kErrInvalidArg;
            for (uint32_t k = s; k < m.patches_offset[p]; ++k)
                if (m.patches_val[k] >= m.nF || m.face_patch[m.patches_val[k]] != p) return kErrInvalidArg;
            s = m.patches_offset[p];
        }
        if (s != m.nF) return kErrInvalidArg;
    } else {
        // faces grouped by patch, ascending face order (patcher.cu:631-637)
        m.patches_offset.resize(P);
        std::vector<uint32_t> off;
        exclusive_scan(cnt, off);
        for (uint32_t p = 0; p < P; ++p) m.patches_offset[p] = off[p + 1];
        m.patches_val.resize(m.nF);
        std::vector<uint32_t> cur(off.begin(), off.end() - 1);
        for (uint32_t f = 0; f < m.nF; ++f) m.patches_val[cur[m.face_patch[f]]++] = f;
    }
    return kOk;
}

// ---------------------------------------------------------------------------
// a3..a6. reference: patcher.cu:640-757 (ribbons, ownership) and
// rxmesh.cpp:754-996 (ltog + topology). Ownership uses the reference rule
// "first patch in patch order with a face touching the element", which is
// the minimum patch id over incident faces.
int build_patches(HostMesh& m)
{
    const uint32_t P = m.P;
    m.vertex_patch.assign(m.nV, kInvalid32);
    m.edge_patch.assign(m.nE, kInvalid32);
    for (uint32_t v = 0; v < m.nV; ++v)
        for (uint32_t j = m.vf_off[v]; j < m.vf_off[v + 1]; ++j)
            m.vertex_patch[v] = std::min(m.vertex_patch[v], m.face_patch[m.vf_val[j]]);
    for (uint32_t e = 0; e < m.nE; ++e)
        for (uint32_t j = m.ef_off[e]; j < m.ef_off[e + 1]; ++j)
            m.edge_patch[e] = std::min(m.edge_patch[e], m.face_patch[m.ef_val[j]]);

    struct PatchLists
    {
        std::vector<uint32_t> ribbon;          // discovery order
        std::vector<uint32_t> lf, le, lv;      // ltog
        uint32_t              nfo, neo, nvo;
    };
    std::vector<PatchLists> pl(P);
    std::atomic<int>        err{kOk};
    parallel_for(P, [&](size_t pi) {
        const uint32_t p  = static_cast<uint32_t>(pi);
        const uint32_t ps = (p == 0) ? 0 : m.patches_offset[p - 1];
        const uint32_t pe = m.patches_offset[p];
        PatchLists&    L  = pl[p];
        // ribbon: faces around the vertices of owned faces, first discovery
        std::vector<uint32_t> cand;
        for (uint32_t k = ps; k < pe; ++k) {
            uint32_t f = m.patches_val[k];
            for (int i = 0; i < 3; ++i) {
                uint32_t v = m.fv[3ull * f + i];
                for (uint32_t j = m.vf_off[v]; j < m.vf_off[v + 1]; ++j)
                    if (m.face_patch[m.vf_val[j]] != p) cand.push_back(m.vf_val[j]);
            }
        }
        {
            std::vector<std::pair<uint32_t, uint32_t>> s(cand.size());
            for (uint32_t i = 0; i < cand.size(); ++i) s[i] = {cand[i], i};
            std::sort(s.begin(), s.end());
            std::vector<uint32_t> firsts;
            for (size_t i = 0; i < s.size(); ++i)
                if (i == 0 || s[i].first != s[i - 1].first) firsts.push_back(s[i].second);
            std::sort(firsts.begin(), firsts.end());
            L.ribbon.clear();
            for (uint32_t i : firsts) L.ribbon.push_back(cand[i]);
        }
        // faces: owned sorted, then ribbon sorted
        L.lf.assign(m.patches_val.begin() + ps, m.patches_val.begin() + pe);
        std::sort(L.lf.begin(), L.lf.end());
        L.nfo = static_cast<uint32_t>(L.lf.size());
        std::vector<uint32_t> rs = L.ribbon;
        std::sort(rs.begin(), rs.end());
        L.lf.insert(L.lf.end(), rs.begin(), rs.end());
        // vertices and edges of all local faces; owned part first, each sorted
        std::vector<uint32_t> vs, es;
        vs.reserve(3 * L.lf.size());
        es.reserve(3 * L.lf.size());
        for (uint32_t f : L.lf)
            for (int i = 0; i < 3; ++i) {
                vs.push_back(m.fv[3ull * f + i]);
                es.push_back(m.fe_g[3ull * f + i]);
            }
        auto uniq_part = [&](std::vector<uint32_t>& x, const std::vector<uint32_t>& owner,
                             uint32_t& n_owned) {
            std::sort(x.begin(), x.end());
            x.erase(std::unique(x.begin(), x.end()), x.end());
            auto it = std::stable_partition(x.begin(), x.end(),
                                            [&](uint32_t g) { return owner[g] == p; });
            n_owned = static_cast<uint32_t>(it - x.begin());
        };
        uniq_part(vs, m.vertex_patch, L.nvo);
        uniq_part(es, m.edge_patch, L.neo);
        L.lv = std::move(vs);
        L.le = std::move(es);
        if (L.lv.size() >= kInvalid16 || L.le.size() >= kInvalid16 || L.lf.size() >= kInvalid16)
            err = kErrPatchTooBig;
    }, 1);
    if (err != kOk) return err;

    // offsets and prefixes
    m.nv.resize(P); m.ne.resize(P); m.nf.resize(P);
    m.nvo.resize(P); m.neo.resize(P); m.nfo.resize(P);
    m.v_off.assign(P + 1, 0); m.e_off.assign(P + 1, 0); m.f_off.assign(P + 1, 0);
    m.vp.assign(P + 1, 0); m.ep.assign(P + 1, 0); m.fp.assign(P + 1, 0);
    m.ribbon_offset.resize(P);
    m.ribbon_val.clear();
    m.max_nv = m.max_ne = m.max_nf = m.max_nvo = m.max_neo = m.max_nfo = 0;
    for (uint32_t p = 0; p < P; ++p) {
        const PatchLists& L = pl[p];
        m.nv[p] = static_cast<uint32_t>(L.lv.size());
        m.ne[p] = static_cast<uint32_t>(L.le.size());
        m.nf[p] = static_cast<uint32_t>(L.lf.size());
        m.nvo[p] = L.nvo; m.neo[p] = L.neo; m.nfo[p] = L.nfo;
        m.v_off[p + 1] = m.v_off[p] + round_up(m.nv[p], kSlabAlign);
        m.e_off[p + 1] = m.e_off[p] + round_up(m.ne[p], kSlabAlign);
        m.f_off[p + 1] = m.f_off[p] + round_up(m.nf[p], kSlabAlign);
        m.vp[p + 1] = m.vp[p] + L.nvo;
        m.ep[p + 1] = m.ep[p] + L.neo;
        m.fp[p + 1] = m.fp[p] + L.nfo;
        m.max_nv = std::max(m.max_nv, m.nv[p]);
        m.max_ne = std::max(m.max_ne, m.ne[p]);
        m.max_nf = std::max(m.max_nf, m.nf[p]);
        m.max_nvo = std::max(m.max_nvo, L.nvo);
        m.max_neo = std::max(m.max_neo, L.neo);
        m.max_nfo = std::max(m.max_nfo, L.nfo);
        m.ribbon_val.insert(m.ribbon_val.end(), L.ribbon.begin(), L.ribbon.end());
        m.ribbon_offset[p] = static_cast<uint32_t>(m.ribbon_val.size());
    }
    if (m.vp[P] != m.nV || m.ep[P] != m.nE || m.fp[P] != m.nF) return kErrInvalidArg;  // isolated vertices

    m.ltog_v.assign(m.v_off[P], kInvalid32);
    m.ltog_e.assign(m.e_off[P], kInvalid32);
    m.ltog_f.assign(m.f_off[P], kInvalid32);
    m.v_l2g.resize(m.nV); m.e_l2g.resize(m.nE); m.f_l2g.resize(m.nF);
    m.v_g2l.resize(m.nV); m.e_g2l.resize(m.nE); m.f_g2l.resize(m.nF);
    parallel_for(P, [&](size_t pi) {
        const uint32_t p = static_cast<uint32_t>(pi);
        const PatchLists& L = pl[p];
        std::copy(L.lv.begin(), L.lv.end(), m.ltog_v.begin() + m.v_off[p]);
        std::copy(L.le.begin(), L.le.end(), m.ltog_e.begin() + m.e_off[p]);
        std::copy(L.lf.begin(), L.lf.end(), m.ltog_f.begin() + m.f_off[p]);
        for (uint32_t l = 0; l < L.nvo; ++l) {
            m.v_l2g[m.vp[p] + l] = L.lv[l];
            m.v_g2l[L.lv[l]]     = m.vp[p] + l;
        }
        for (uint32_t l = 0; l < L.neo; ++l) {
            m.e_l2g[m.ep[p] + l] = L.le[l];
            m.e_g2l[L.le[l]]     = m.ep[p] + l;
        }
        for (uint32_t l = 0; l < L.nfo; ++l) {
            m.f_l2g[m.fp[p] + l] = L.lf[l];
            m.f_g2l[L.lf[l]]     = m.fp[p] + l;
        }
    }, 1);

    // topology and owner tables
    m.ev_local.assign(m.e_off[P], 0xFFFFFFFFu);
    m.fe_local.assign(3ull * m.f_off[P], kInvalid16);
    m.v_lin.assign(m.v_off[P], kInvalid32); m.v_own.assign(m.v_off[P], kInvalid32);
    m.e_lin.assign(m.e_off[P], kInvalid32); m.e_own.assign(m.e_off[P], kInvalid32);
    m.f_lin.assign(m.f_off[P], kInvalid32); m.f_own.assign(m.f_off[P], kInvalid32);
    parallel_for(P, [&](size_t pi) {
        const uint32_t p = static_cast<uint32_t>(pi);
        const PatchLists& L = pl[p];
        auto local_of = [&](const std::vector<uint32_t>& l, uint32_t nown, bool owned,
                            uint32_t g) -> uint32_t {
            auto b = owned ? l.begin() : l.begin() + nown;
            auto e = owned ? l.begin() + nown : l.end();
            auto it = std::lower_bound(b, e, g);
            return static_cast<uint32_t>(it - l.begin());
        };
        for (uint32_t le = 0; le < L.le.size(); ++le) {
            uint32_t g = L.le[le];
            uint32_t a = m.ev[2ull * g], b = m.ev[2ull * g + 1];
            uint32_t hi = std::max(a, b), lo = std::min(a, b);
            uint32_t l0 = local_of(L.lv, L.nvo, m.vertex_patch[hi] == p, hi);
            uint32_t l1 = local_of(L.lv, L.nvo, m.vertex_patch[lo] == p, lo);
            m.ev_local[m.e_off[p] + le] = l0 | (l1 << 16);
        }
        for (uint32_t lf = 0; lf < L.lf.size(); ++lf) {
            uint32_t g = L.lf[lf];
            for (int i = 0; i < 3; ++i) {
                uint32_t a  = m.fv[3ull * g + i], b = m.fv[3ull * g + (i + 1) % 3];
                uint32_t ge = m.fe_g[3ull * g + i];
                uint32_t le = local_of(L.le, L.neo, m.edge_patch[ge] == p, ge);
                uint32_t dir = (a == std::max(a, b)) ? 0u : 1u;
                m.fe_local[3ull * m.f_off[p] + 3ull * lf + i] = static_cast<uint16_t>((le << 1) | dir);
            }
        }
        for (uint32_t l = 0; l < L.lv.size(); ++l) {
            m.v_lin[m.v_off[p] + l] = m.v_g2l[L.lv[l]];
            m.v_own[m.v_off[p] + l] = m.vertex_patch[L.lv[l]];
        }
        for (uint32_t l = 0; l < L.le.size(); ++l) {
            m.e_lin[m.e_off[p] + l] = m.e_g2l[L.le[l]];
            m.e_own[m.e_off[p] + l] = m.edge_patch[L.le[l]];
        }
        for (uint32_t l = 0; l < L.lf.size(); ++l) {
            m.f_lin[m.f_off[p] + l] = m.f_g2l[L.lf[l]];
            m.f_own[m.f_off[p] + l] = m.face_patch[L.lf[l]];
        }
    }, 1);
    return kOk;
}

// ---------------------------------------------------------------------------
// Closed-fan vertices (host_build.h).
void closed_fan_vertices(const HostMesh& m, std::vector<uint8_t>& closed)
{
    std::vector<uint32_t> valence(m.nV, 0);
    for (uint32_t e = 0; e < m.nE; ++e) {
        valence[m.ev[2ull * e]]++;
        valence[m.ev[2ull * e + 1]]++;
    }
    closed.assign(m.nV, 0);
    parallel_for(m.nV, [&](size_t v) {
        const uint32_t b = m.vf_off[v], k = m.vf_off[v + 1] - b;
        if (k == 0 || k != valence[v]) return;
        std::vector<uint32_t> key(k), nxt(k);
        for (uint32_t i = 0; i < k; ++i) {
            const uint32_t f = m.vf_val[b + i];
            uint32_t       j = 0;
            while (j < 3 && m.fv[3ull * f + j] != v) ++j;
            if (j == 3) return;
            key[i] = m.fv[3ull * f + (j + 1) % 3];
            nxt[i] = m.fv[3ull * f + (j + 2) % 3];
        }
        // follow the successor map from key[0]; a single cycle visits k keys
        uint32_t cur = 0, steps = 0;
        do {
            uint32_t i = 0;
            while (i < k && key[i] != nxt[cur]) ++i;
            if (i == k) return;
            cur = i;
            ++steps;
        } while (cur != 0 && steps <= k);
        if (cur == 0 && steps == k) closed[v] = 1;
    }, 4096);
}

// ---------------------------------------------------------------------------
// Patch-local VV edge pattern (host_build.h).
bool build_vv_edge_pattern(const HostMesh& m, const std::vector<uint32_t>& vv_base, std::vector<uint16_t>& ent,
                           std::vector<uint32_t>& rs)
{
    for (uint32_t p = 0; p < m.P; ++p)
        if (m.ne[p] >= 0x8000u) return false;
    ent.assign(vv_base[m.P], 0);
    rs.assign(m.nV + 1, 0);
    rs[m.nV] = vv_base[m.P];
    parallel_for(m.P, [&](size_t p) {
        const uint32_t        nvo = m.nvo[p], ne = m.ne[p], vp = m.vp[p];
        std::vector<uint32_t> cnt(nvo + 1, 0);
        const uint32_t*       ev = m.ev_local.data() + m.e_off[p];
        for (uint32_t e = 0; e < ne; ++e) {
            const uint32_t a = ev[e] & 0xFFFFu, b = ev[e] >> 16;
            if (a < nvo) cnt[a + 1]++;
            if (b < nvo) cnt[b + 1]++;
        }
        for (uint32_t l = 0; l < nvo; ++l) cnt[l + 1] += cnt[l];
        for (uint32_t l = 0; l < nvo; ++l) rs[vp + l] = vv_base[p] + cnt[l];
        uint16_t* out = ent.data() + vv_base[p];
        for (uint32_t e = 0; e < ne; ++e) {
            const uint32_t a = ev[e] & 0xFFFFu, b = ev[e] >> 16;
            if (a < nvo) out[cnt[a]++] = static_cast<uint16_t>(2u * e);
            if (b < nvo) out[cnt[b]++] = static_cast<uint16_t>(2u * e + 1u);
        }
    }, 16);
    return true;
}

// ---------------------------------------------------------------------------
// Two-step smoothing plan (host_build.h).
int build_vv2_plan(const HostMesh& m, Vv2Plan& out)
{
    const uint32_t P = m.P, V = m.nV;
    // canonical VV rows in global ids, indexed by linear vertex id: an owned
    // vertex's row lists the other endpoints of its incident edges in
    // ascending owner-local edge order
    std::vector<uint32_t> deg(V, 0);
    for (uint64_t e = 0; e < m.nE; ++e) {
        deg[m.ev[2 * e]]++;
        deg[m.ev[2 * e + 1]]++;
    }
    std::vector<uint32_t> roff_g(V + 1, 0);
    for (uint32_t l = 0; l < V; ++l) roff_g[l + 1] = roff_g[l] + deg[m.v_l2g[l]];
    std::vector<uint32_t> rows(roff_g[V]);
    parallel_for(P, [&](size_t q) {
        const uint32_t        vo = m.v_off[q], eo = m.e_off[q], nvo = m.nvo[q], ne = m.ne[q];
        std::vector<uint32_t> fill(nvo, 0);
        for (uint32_t e = 0; e < ne; ++e) {
            const uint32_t x = m.ev_local[eo + e], a = x & 0xFFFFu, b = x >> 16;
            if (a < nvo) rows[roff_g[m.vp[q] + a] + fill[a]++] = m.ltog_v[vo + b];
            if (b < nvo) rows[roff_g[m.vp[q] + b] + fill[b]++] = m.ltog_v[vo + a];
        }
    }, 1);
    // per patch: extended numbering, rows of all local vertices, extras
    std::vector<std::vector<uint16_t>> pr(P), pn(P);
    std::vector<std::vector<uint32_t>> px(P);
    std::atomic<int>                   err{kOk};
    parallel_for(P, [&](size_t p) {
        const uint32_t  vo = m.v_off[p], nv = m.nv[p], nvo = m.nvo[p];
        const uint32_t* gl = &m.ltog_v[vo];  // owned part sorted, ribbon part sorted
        auto local_of = [&](uint32_t g) -> uint32_t {
            const uint32_t* a = std::lower_bound(gl, gl + nvo, g);
            if (a != gl + nvo && *a == g) return static_cast<uint32_t>(a - gl);
            const uint32_t* b = std::lower_bound(gl + nvo, gl + nv, g);
            if (b != gl + nv && *b == g) return static_cast<uint32_t>(b - gl);
            return kInvalid32;
        };
        auto row_of = [&](uint32_t l) {
            const uint32_t lin = (l < nvo) ? m.vp[p] + l : m.v_lin[vo + l];
            return std::make_pair(roff_g[lin], roff_g[lin + 1]);
        };
        std::vector<uint32_t>& X = px[p];
        for (uint32_t l = 0; l < nv; ++l) {
            const auto r = row_of(l);
            for (uint32_t k = r.first; k < r.second; ++k)
                if (local_of(rows[k]) == kInvalid32) X.push_back(rows[k]);
        }
        std::sort(X.begin(), X.end());
        X.erase(std::unique(X.begin(), X.end()), X.end());
        std::vector<uint16_t>& R = pr[p];
        std::vector<uint16_t>& Nb = pn[p];
        R.assign(nv + 1, 0);
        uint64_t total = 0;
        for (uint32_t l = 0; l < nv; ++l) {
            const auto r = row_of(l);
            total += r.second - r.first;
        }
        if (nv + X.size() >= 0xFFFFu || total >= 0xFFFFu) {
            err = kErrPatchTooBig;
            return;
        }
        Nb.reserve(total);
        for (uint32_t l = 0; l < nv; ++l) {
            const auto r = row_of(l);
            for (uint32_t k = r.first; k < r.second; ++k) {
                uint32_t t = local_of(rows[k]);
                if (t == kInvalid32)
                    t = nv + static_cast<uint32_t>(std::lower_bound(X.begin(), X.end(), rows[k]) - X.begin());
                Nb.push_back(static_cast<uint16_t>(t));
            }
            R[l + 1] = static_cast<uint16_t>(Nb.size());
        }
        for (uint32_t& g : X) g = m.v_g2l[g];  // gather list: linear ids
    });
    if (err != kOk) return err;
    out.rbase.assign(P + 1, 0);
    out.nbase.assign(P + 1, 0);
    out.xbase.assign(P + 1, 0);
    out.nx.assign(P, 0);
    out.max_nx = out.max_nent = 0;
    for (uint32_t p = 0; p < P; ++p) {
        out.rbase[p + 1] = out.rbase[p] + round_up(static_cast<uint32_t>(pr[p].size()), 8);
        out.nbase[p + 1] = out.nbase[p] + round_up(static_cast<uint32_t>(pn[p].size()), 8);
        out.xbase[p + 1] = out.xbase[p] + round_up(static_cast<uint32_t>(px[p].size()), 4);
        out.nx[p]        = static_cast<uint32_t>(px[p].size());
        out.max_nx       = std::max(out.max_nx, out.nx[p]);
        out.max_nent     = std::max(out.max_nent, static_cast<uint32_t>(pn[p].size()));
    }
    out.roff.assign(std::max<uint32_t>(out.rbase[P], 8), 0);
    out.nbr.assign(std::max<uint32_t>(out.nbase[P], 8), 0);
    out.xlin.assign(std::max<uint32_t>(out.xbase[P], 4), 0);
    parallel_for(P, [&](size_t p) {
        std::copy(pr[p].begin(), pr[p].end(), out.roff.begin() + out.rbase[p]);
        std::copy(pn[p].begin(), pn[p].end(), out.nbr.begin() + out.nbase[p]);
        std::copy(px[p].begin(), px[p].end(), out.xlin.begin() + out.xbase[p]);
    }, 64);
    return kOk;
}

// K-level smoothing plan (host_build.h).
int build_vvk_plan(const HostMesh& m, uint32_t K, VvkPlan& out)
{
    if (K < 1 || K > 16) return kErrInvalidArg;
    const uint32_t P = m.P, V = m.nV, W = 2 * K + 1;
    // canonical VV rows in global ids, indexed by linear vertex id (as in
    // build_vv2_plan): ascending owner-local edge order
    std::vector<uint32_t> deg(V, 0);
    for (uint64_t e = 0; e < m.nE; ++e) {
        deg[m.ev[2 * e]]++;
        deg[m.ev[2 * e + 1]]++;
    }
    std::vector<uint32_t> roff_g(V + 1, 0);
    for (uint32_t l = 0; l < V; ++l) roff_g[l + 1] = roff_g[l] + deg[m.v_l2g[l]];
    std::vector<uint32_t> rows(roff_g[V]);
    parallel_for(P, [&](size_t q) {
        const uint32_t        vo = m.v_off[q], eo = m.e_off[q], nvo = m.nvo[q], ne = m.ne[q];
        std::vector<uint32_t> fill(nvo, 0);
        for (uint32_t e = 0; e < ne; ++e) {
            const uint32_t x = m.ev_local[eo + e], a = x & 0xFFFFu, b = x >> 16;
            if (a < nvo) rows[roff_g[m.vp[q] + a] + fill[a]++] = m.ltog_v[vo + b];
            if (b < nvo) rows[roff_g[m.vp[q] + b] + fill[b]++] = m.ltog_v[vo + a];
        }
    }, 1);
    std::vector<std::vector<uint16_t>> pr(P), pn(P);
    std::vector<std::vector<uint32_t>> px(P);
    out.K = K;
    out.lvl.assign(size_t(P) * W, 0);
    std::atomic<int> err{kOk};
    parallel_for(P, [&](size_t p) {
        const uint32_t  vo = m.v_off[p], nvo = m.nvo[p];
        const uint32_t* gl = &m.ltog_v[vo];
        std::unordered_map<uint32_t, uint32_t> ext;  // global id -> extended id
        std::vector<uint32_t>                  E(gl, gl + nvo);
        ext.reserve(8 * nvo + 64);
        for (uint32_t i = 0; i < nvo; ++i) ext.emplace(gl[i], i);
        uint32_t* c = &out.lvl[size_t(p) * W];
        c[0]        = nvo;
        uint32_t lo = 0;
        for (uint32_t k = 1; k <= K; ++k) {
            std::vector<uint32_t> next;
            for (uint32_t i = lo; i < c[k - 1]; ++i) {
                const uint32_t lin = m.v_g2l[E[i]];
                for (uint32_t j = roff_g[lin]; j < roff_g[lin + 1]; ++j)
                    if (!ext.count(rows[j])) next.push_back(rows[j]);
            }
            std::sort(next.begin(), next.end());
            next.erase(std::unique(next.begin(), next.end()), next.end());
            for (uint32_t g : next) {
                ext.emplace(g, static_cast<uint32_t>(E.size()));
                E.push_back(g);
            }
            lo   = c[k - 1];
            c[k] = static_cast<uint32_t>(E.size());
        }
        if (c[K] >= 0xFFFFu) {
            err = kErrPatchTooBig;
            return;
        }
        std::vector<uint16_t>& R  = pr[p];
        std::vector<uint16_t>& Nb = pn[p];
        const uint32_t         nr = c[K - 1];
        R.assign(nr + 1, 0);
        for (uint32_t i = 0; i < nr; ++i) {
            const uint32_t lin = m.v_g2l[E[i]];
            for (uint32_t j = roff_g[lin]; j < roff_g[lin + 1]; ++j) Nb.push_back(static_cast<uint16_t>(ext.at(rows[j])));
            if (Nb.size() >= 0xFFFFu) {
                err = kErrPatchTooBig;
                return;
            }
            R[i + 1] = static_cast<uint16_t>(Nb.size());
        }
        for (uint32_t k = 0; k < K; ++k) c[K + 1 + k] = R[c[k]];
        std::vector<uint32_t>& X = px[p];
        X.resize(c[K] - nvo);
        for (uint32_t i = nvo; i < c[K]; ++i) X[i - nvo] = m.v_g2l[E[i]];
    });
    if (err != kOk) return err;
    out.rbase.assign(P + 1, 0);
    out.nbase.assign(P + 1, 0);
    out.xbase.assign(P + 1, 0);
    out.max_cnt.assign(K + 1, 0);
    out.max_ent.assign(K, 0);
    for (uint32_t p = 0; p < P; ++p) {
        out.rbase[p + 1] = out.rbase[p] + round_up(static_cast<uint32_t>(pr[p].size()), 8);
        out.nbase[p + 1] = out.nbase[p] + round_up(static_cast<uint32_t>(pn[p].size()), 8);
        out.xbase[p + 1] = out.xbase[p] + round_up(static_cast<uint32_t>(px[p].size()), 4);
        const uint32_t* c = &out.lvl[size_t(p) * W];
        for (uint32_t k = 0; k <= K; ++k) out.max_cnt[k] = std::max(out.max_cnt[k], c[k]);
        for (uint32_t k = 0; k < K; ++k) out.max_ent[k] = std::max(out.max_ent[k], c[K + 1 + k]);
    }
    out.roff.assign(std::max<uint32_t>(out.rbase[P], 8), 0);
    out.nbr.assign(std::max<uint32_t>(out.nbase[P], 8), 0);
    out.xlin.assign(std::max<uint32_t>(out.xbase[P], 4), 0);
    parallel_for(P, [&](size_t p) {
        std::copy(pr[p].begin(), pr[p].end(), out.roff.begin() + out.rbase[p]);
        std::copy(pn[p].begin(), pn[p].end(), out.nbr.begin() + out.nbase[p]);
        std::copy(px[p].begin(), px[p].end(), out.xlin.begin() + out.xbase[p]);
    }, 64);
    return kOk;
}

int concat_meshes(const std::vector<const HostMesh*>& parts, HostMesh& o, std::vector<uint32_t>& vb,
                  std::vector<uint32_t>& eb, std::vector<uint32_t>& fb, std::vector<uint32_t>& pb)
{
    const size_t K = parts.size();
    if (K == 0) return kErrInvalidArg;
    vb.assign(K + 1, 0); eb.assign(K + 1, 0); fb.assign(K + 1, 0); pb.assign(K + 1, 0);
    std::vector<uint32_t> vs(K + 1, 0), es(K + 1, 0), fs(K + 1, 0);  // slab bases
    std::vector<uint64_t> hb(K + 1, 0), rb(K + 1, 0);                  // half-edge, ribbon bases
    for (size_t k = 0; k < K; ++k) {
        const HostMesh& m = *parts[k];
        uint64_t nv = uint64_t(vb[k]) + m.nV, ne = uint64_t(eb[k]) + m.nE, nf = uint64_t(fb[k]) + m.nF;
        uint64_t np = uint64_t(pb[k]) + m.P;
        uint64_t sv = uint64_t(vs[k]) + m.v_off[m.P], se = uint64_t(es[k]) + m.e_off[m.P];
        uint64_t sf = uint64_t(fs[k]) + m.f_off[m.P];
        if (nv >= kInvalid32 || ne >= kInvalid32 || 3 * nf >= kInvalid32 || np >= kInvalid32 ||
            sv >= kInvalid32 || se >= kInvalid32 || 3 * sf >= kInvalid32)
            return kErrInvalidArg;
        vb[k + 1] = uint32_t(nv); eb[k + 1] = uint32_t(ne); fb[k + 1] = uint32_t(nf); pb[k + 1] = uint32_t(np);
        vs[k + 1] = uint32_t(sv); es[k + 1] = uint32_t(se); fs[k + 1] = uint32_t(sf);
        hb[k + 1] = hb[k] + 3ull * m.nF;
        rb[k + 1] = rb[k] + m.ribbon_val.size();
    }
    o = HostMesh();
    o.nV = vb[K]; o.nE = eb[K]; o.nF = fb[K]; o.P = pb[K];
    o.patch_size = parts[0]->patch_size;
    o.closed = true; o.manifold = true; o.oriented = true;
    // the output arrays (several GB for large batches) are sized on all host
    // threads at once, one array per thread: the zeroing and the first-touch
    // page faults of a serial resize would dominate the concatenation
    std::vector<std::pair<std::vector<uint32_t>*, uint64_t>> sz = {
        {&o.fv, 3ull * o.nF},      {&o.fe_g, 3ull * o.nF},    {&o.ev, 2ull * o.nE},      {&o.vf_off, o.nV + 1ull},
        {&o.vf_val, 3ull * o.nF},  {&o.ef_off, o.nE + 1ull},  {&o.ef_val, 3ull * o.nF},  {&o.face_patch, o.nF},
        {&o.patches_val, o.nF},    {&o.patches_offset, o.P},  {&o.ribbon_val, rb[K]},    {&o.ribbon_offset, o.P},
        {&o.vertex_patch, o.nV},   {&o.edge_patch, o.nE},     {&o.ltog_v, vs[K]},        {&o.v_lin, vs[K]},
        {&o.v_own, vs[K]},         {&o.ltog_e, es[K]},        {&o.e_lin, es[K]},         {&o.e_own, es[K]},
        {&o.ev_local, es[K]},      {&o.ltog_f, fs[K]},        {&o.f_lin, fs[K]},         {&o.f_own, fs[K]},
        {&o.v_l2g, o.nV},          {&o.v_g2l, o.nV},          {&o.e_l2g, o.nE},          {&o.e_g2l, o.nE},
        {&o.f_l2g, o.nF},          {&o.f_g2l, o.nF}};
    for (auto* v : {&o.nv, &o.ne, &o.nf, &o.nvo, &o.neo, &o.nfo}) sz.push_back({v, o.P});
    for (auto* v : {&o.v_off, &o.e_off, &o.f_off, &o.vp, &o.ep, &o.fp}) sz.push_back({v, o.P + 1ull});
    parallel_for(sz.size() + 1, [&](size_t i) {
        if (i < sz.size()) sz[i].first->assign(sz[i].second, 0u);
        else o.fe_local.assign(3ull * fs[K], uint16_t(0));
    }, 1);
    for (size_t k = 0; k < K; ++k) {
        const HostMesh& m = *parts[k];
        o.closed = o.closed && m.closed;
        o.manifold = o.manifold && m.manifold;
        o.oriented = o.oriented && m.oriented;
        o.max_valence = std::max(o.max_valence, m.max_valence);
        o.max_edge_faces = std::max(o.max_edge_faces, m.max_edge_faces);
        o.max_face_adj = std::max(o.max_face_adj, m.max_face_adj);
        o.max_nv = std::max(o.max_nv, m.max_nv); o.max_ne = std::max(o.max_ne, m.max_ne);
        o.max_nf = std::max(o.max_nf, m.max_nf); o.max_nvo = std::max(o.max_nvo, m.max_nvo);
        o.max_neo = std::max(o.max_neo, m.max_neo); o.max_nfo = std::max(o.max_nfo, m.max_nfo);
    }
    parallel_for(K, [&](size_t k) {
        const HostMesh& m = *parts[k];
        const uint32_t V0 = vb[k], E0 = eb[k], F0 = fb[k], P0 = pb[k];
        for (uint64_t i = 0; i < 3ull * m.nF; ++i) {
            o.fv[3ull * F0 + i]     = m.fv[i] + V0;
            o.fe_g[3ull * F0 + i]   = m.fe_g[i] + E0;
            o.vf_val[3ull * F0 + i] = m.vf_val[i] + F0;
            o.ef_val[3ull * F0 + i] = m.ef_val[i] + F0;
        }
        for (uint64_t i = 0; i < 2ull * m.nE; ++i) o.ev[2ull * E0 + i] = m.ev[i] + V0;
        for (uint32_t v = 0; v < m.nV; ++v) {
            o.vf_off[V0 + v + 1] = m.vf_off[v + 1] + uint32_t(hb[k]);
            o.vertex_patch[V0 + v] = m.vertex_patch[v] + P0;
            o.v_l2g[V0 + v] = m.v_l2g[v] + V0;
            o.v_g2l[V0 + v] = m.v_g2l[v] + V0;
        }
        for (uint32_t e = 0; e < m.nE; ++e) {
            o.ef_off[E0 + e + 1] = m.ef_off[e + 1] + uint32_t(hb[k]);
            o.edge_patch[E0 + e] = m.edge_patch[e] + P0;
            o.e_l2g[E0 + e] = m.e_l2g[e] + E0;
            o.e_g2l[E0 + e] = m.e_g2l[e] + E0;
        }
        for (uint32_t f = 0; f < m.nF; ++f) {
            o.face_patch[F0 + f] = m.face_patch[f] + P0;
            o.patches_val[F0 + f] = m.patches_val[f] + F0;
            o.f_l2g[F0 + f] = m.f_l2g[f] + F0;
            o.f_g2l[F0 + f] = m.f_g2l[f] + F0;
        }
        for (size_t i = 0; i < m.ribbon_val.size(); ++i) o.ribbon_val[rb[k] + i] = m.ribbon_val[i] + F0;
        for (uint32_t p = 0; p < m.P; ++p) {
            const uint32_t q = P0 + p;
            o.patches_offset[q] = m.patches_offset[p] + F0;
            o.ribbon_offset[q] = m.ribbon_offset[p] + uint32_t(rb[k]);
            o.nv[q] = m.nv[p]; o.ne[q] = m.ne[p]; o.nf[q] = m.nf[p];
            o.nvo[q] = m.nvo[p]; o.neo[q] = m.neo[p]; o.nfo[q] = m.nfo[p];
            o.v_off[q] = m.v_off[p] + vs[k]; o.e_off[q] = m.e_off[p] + es[k]; o.f_off[q] = m.f_off[p] + fs[k];
            o.vp[q] = m.vp[p] + V0; o.ep[q] = m.ep[p] + E0; o.fp[q] = m.fp[p] + F0;
        }
        for (uint32_t i = 0; i < m.v_off[m.P]; ++i) {
            const bool pad = m.ltog_v[i] == kInvalid32;
            o.ltog_v[vs[k] + i] = pad ? kInvalid32 : m.ltog_v[i] + V0;
            o.v_lin[vs[k] + i] = pad ? kInvalid32 : m.v_lin[i] + V0;
            o.v_own[vs[k] + i] = pad ? kInvalid32 : m.v_own[i] + P0;
        }
        for (uint32_t i = 0; i < m.e_off[m.P]; ++i) {
            const bool pad = m.ltog_e[i] == kInvalid32;
            o.ltog_e[es[k] + i] = pad ? kInvalid32 : m.ltog_e[i] + E0;
            o.e_lin[es[k] + i] = pad ? kInvalid32 : m.e_lin[i] + E0;
            o.e_own[es[k] + i] = pad ? kInvalid32 : m.e_own[i] + P0;
            o.ev_local[es[k] + i] = m.ev_local[i];
        }
        for (uint32_t i = 0; i < m.f_off[m.P]; ++i) {
            const bool pad = m.ltog_f[i] == kInvalid32;
            o.ltog_f[fs[k] + i] = pad ? kInvalid32 : m.ltog_f[i] + F0;
            o.f_lin[fs[k] + i] = pad ? kInvalid32 : m.f_lin[i] + F0;
            o.f_own[fs[k] + i] = pad ? kInvalid32 : m.f_own[i] + P0;
        }
        for (uint64_t i = 0; i < 3ull * m.f_off[m.P]; ++i) o.fe_local[3ull * fs[k] + i] = m.fe_local[i];
    }, 1);
    o.v_off[o.P] = vs[K]; o.e_off[o.P] = es[K]; o.f_off[o.P] = fs[K];
    o.vp[o.P] = o.nV; o.ep[o.P] = o.nE; o.fp[o.P] = o.nF;
    return kOk;
}

// ---------------------------------------------------------------------------
// Reference patch archive: cereal PortableBinary of Patcher::serialize
// (include/rxmesh/patcher/patcher.h:162-182): u8 little-endian flag, nine u32
// scalars, seven (u64 size + u32[size]) vectors, one f32.
int read_patch_file(const std::string& path, std::vector<uint32_t>& face_patch,
                    std::vector<uint32_t>& patches_val, std::vector<uint32_t>& patches_offset,
                    uint32_t& num_patches, uint32_t& nV, uint32_t& nE, uint32_t& nF)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) return kErrIO;
    uint8_t le = 0;
    in.read(reinterpret_cast<char*>(&le), 1);
    if (!in || le != 1) return kErrUnsupported;
    uint32_t hdr[9];
    in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    if (!in) return kErrIO;
    num_patches = hdr[1];
    nV = hdr[2];
    nE = hdr[3];
    nF = hdr[4];
    std::vector<uint32_t> vecs[7];
    for (auto& v : vecs) {
        uint64_t n = 0;
        in.read(reinterpret_cast<char*>(&n), 8);
        if (!in || n > (1ull << 32)) return kErrIO;
        v.resize(n);
        in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(4 * n));
        if (!in) return kErrIO;
    }
    face_patch     = std::move(vecs[0]);
    patches_val    = std::move(vecs[3]);
    patches_offset = std::move(vecs[4]);
    if (face_patch.size() != nF || patches_val.size() != nF || patches_offset.size() < num_patches)
        return kErrIO;
    patches_offset.resize(num_patches);
    return kOk;
}

int write_patch_file(const std::string& path, const HostMesh& m)
{
    std::ofstream out(path, std::ios::binary);
    if (!out) return kErrIO;
    uint8_t le = 1;
    out.write(reinterpret_cast<const char*>(&le), 1);
    uint32_t hdr[9] = {m.patch_size, m.P, m.nV, m.nE, m.nF, m.P, m.P, 1, 0};
    out.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    const std::vector<uint32_t>* vecs[7] = {&m.face_patch, &m.vertex_patch, &m.edge_patch,
                                            &m.patches_val, &m.patches_offset, &m.ribbon_val,
                                            &m.ribbon_offset};
    for (auto* v : vecs) {
        uint64_t n = v->size();
        out.write(reinterpret_cast<const char*>(&n), 8);
        out.write(reinterpret_cast<const char*>(v->data()), static_cast<std::streamsize>(4 * n));
    }
    float t = 0.f;
    out.write(reinterpret_cast<const char*>(&t), 4);
    return out ? kOk : kErrIO;
}

}  // namespace rxm
