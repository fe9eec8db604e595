// dgx_serve.cpp -- request server for streams of Poisson problems on point
// clouds (one device, continuous admission).
//
// Each request carries a point cloud; the server builds its octree
// (points2Octree + balanceOctree), its multigrid hierarchy (DAMGCreateAndSetDA)
// and solves -lap u + u = f with homogeneous Dirichlet conditions by
// MG-preconditioned CG (DAMGSolve), f nodal N(0,1) from the request's seed.
// Workers ("slots") each own a HIP stream at default priority; a request is
// admitted as soon as a slot is free and its memory estimate fits the device
// budget next to the running requests and the resident problems, so requests
// that arrive while others run start at once and their kernels overlap on the
// GPU.  Finished problems stay resident (octree, DA maps, hierarchy) in an LRU
// cache keyed by a hash of the request, and a repeated request only re-solves.
// Resident problems are evicted, oldest first, when admission needs room.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/dendro_mi355x.h"
#include "dgx_mg.h"

bool dgx_ctx_device(int *device);

namespace {

constexpr int SB = 256;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// f_i = N(0,1) by Box-Muller from a counter-based SplitMix64 stream.
__global__ void k_gen_normal(uint64_t seed, uint32_t n, double *f) {
    const uint32_t i = blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    const uint64_t base = seed * 0xD1B54A32D192ED03ull;
    const double u1 = 1.0 - (double)(mix64(base + (2ull * i + 1) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1p-53;
    const double u2 = (double)(mix64(base + (2ull * i + 2) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1p-53;
    f[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

uint64_t hash_request(const double *xyz, size_t n, unsigned md, unsigned m) {
    // four independent multiply-xorshift lanes over the 64-bit words
    const uint64_t *w = reinterpret_cast<const uint64_t *>(xyz);
    const size_t nw = 3 * n;
    uint64_t h[4] = {0x9E3779B97F4A7C15ull ^ n, 0xC2B2AE3D27D4EB4Full ^ md, 0x165667B19E3779F9ull ^ m,
                     0x27D4EB2F165667C5ull};
    size_t i = 0;
    for (; i + 4 <= nw; i += 4)
        for (int k = 0; k < 4; ++k) {
            h[k] = (h[k] ^ w[i + k]) * 0xBF58476D1CE4E5B9ull;
            h[k] ^= h[k] >> 31;
        }
    for (; i < nw; ++i) h[0] = ((h[0] ^ w[i]) * 0x94D049BB133111EBull) ^ (h[0] >> 29);
    uint64_t r = h[0];
    for (int k = 1; k < 4; ++k) r = (r ^ (h[k] + 0x9E3779B97F4A7C15ull + (r << 6) + (r >> 2))) * 0xBF58476D1CE4E5B9ull;
    return r ^ (r >> 31);
}

struct Request {
    uint64_t id = 0, key = 0, seed = 0, est = 0, npts = 0;
    std::vector<double> xyz;
    unsigned md = 30, mnp = 1;
    double t_submit = 0;
};

struct Resident {
    uint64_t key = 0, bytes = 0;
    std::unique_ptr<MgHierarchy> mg;
    uint64_t octants = 0;
    bool busy = false;
};

struct Done {
    dgx_served r{};
    std::string err;
    std::vector<double> x;
};

}  // namespace

struct dgx_server_s {
    dgx_server_opts o{};
    int device = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> queue;
    std::unordered_map<uint64_t, Done> done;
    std::unordered_set<uint64_t> consumed;  // ids whose result dgx_server_wait has handed out
    std::list<Resident> cache;  // front = most recently used
    std::list<Resident> evicted;  // taken out of the cache under mu, freed by the evicting worker after it
    uint64_t reserved = 0;      // estimates of running requests + bytes of resident problems
    uint64_t next_id = 1;
    uint64_t running = 0, evictions = 0;
    int fail_next_resolve = 0;  // fault injection for tests: the next resident re-solve fails
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    // Bytes a request may hold on the device (points, keys, octree, DA maps,
    // hierarchy and their temporaries scale with the number of octants, i.e.
    // with the number of points over maxNumPts): the largest peak per point
    // seen so far for the request's maxNumPts class (1, 2-4, more), starting
    // from 14000 / 6000 / 3000 B, plus 64 MiB.
    double bytes_per_point[3] = {14000.0, 6000.0, 3000.0};  // under mu
    static int mclass(unsigned m) { return m <= 1 ? 0 : (m <= 4 ? 1 : 2); }
    uint64_t estimate(size_t n, unsigned m) const {
        return (uint64_t)((double)n * bytes_per_point[mclass(m)]) + (64ull << 20);
    }

    // Under mu: make room for `need` by evicting idle resident problems,
    // least recently used first.
    bool fits(uint64_t need) {
        while (reserved + need > o.hbm_budget) {
            auto victim = cache.end();
            for (auto it = cache.begin(); it != cache.end(); ++it)
                if (!it->busy) victim = it;  // last idle entry = least recently used
            if (victim == cache.end()) return false;
            reserved -= victim->bytes;
            // freed outside the lock by the worker that evicts (its stream)
            evicted.splice(evicted.end(), cache, victim);
            ++evictions;
        }
        return true;
    }

    void worker(unsigned slot) {
        (void)slot;
        (void)hipSetDevice(device);
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return;
        // Resident problems evicted by this thread were built on other slots'
        // streams, which may be capturing a graph right now: their (finished)
        // memory is returned on this thread's stream instead.
        dgx_alloc_free_stream(s);
        std::list<Resident> dead;
        for (;;) {
            Request rq;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    if (queue.empty()) return false;
                    // a request larger than the whole budget still runs once nothing else holds memory
                    return fits(queue.front().est) || reserved == 0;
                });
                if (stop && queue.empty()) break;
                rq = std::move(queue.front());
                queue.pop_front();
                reserved += rq.est;
                ++running;
                dead.splice(dead.end(), evicted);
            }
            // the problems this admission evicted: their memory goes back here,
            // without the lock, before this request allocates
            dead.clear();
            Done d;
            d.r.id = rq.id;
            d.r.t_submit = rq.t_submit;
            d.r.t_start = now();
            uint64_t keep_bytes = 0;
            // device memory of this request is counted under its id
            const uint64_t tag = (1ull << 63) | rq.id;
            dgx_alloc_set_tag(tag);
            try {
                run(rq, s, d, keep_bytes, tag);
            } catch (DgxError &) {
                d.r.status = -1;
                d.err = dgx_last_error();
            } catch (std::exception &e) {
                d.r.status = -1;
                d.err = e.what();
            }
            dgx_alloc_set_tag(0);
            (void)hipStreamSynchronize(s);
            d.r.t_done = now();
            d.r.resident_bytes = keep_bytes;
            {
                std::lock_guard<std::mutex> lk(mu);
                reserved -= rq.est;
                reserved += keep_bytes;
                --running;
                if (rq.npts >= 65536 && d.r.peak_bytes > 0)  // small ones are covered by the 64 MiB term
                    bytes_per_point[mclass(rq.mnp)] =
                        std::max(bytes_per_point[mclass(rq.mnp)], (double)d.r.peak_bytes / (double)rq.npts);
                done[rq.id] = std::move(d);
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
        (void)hipStreamSynchronize(s);
        dgx_alloc_free_stream(nullptr);
        std::lock_guard<std::mutex> lk(mu);
        streams.push_back(s);  // destroyed by the server after the resident problems
    }

    // A resident problem taken by a request: on every exit path the request's
    // stream is drained and the hierarchy marked idle, so another slot can
    // solve it again; when the request failed, the problem is dropped from
    // the cache and its bytes are returned to the budget instead, since its
    // state may be partly overwritten.
    struct ResidentLease {
        dgx_server_s *sv = nullptr;
        Resident *res = nullptr;
        hipStream_t s = nullptr;
        bool ok = false;
        ~ResidentLease() {
            if (!res) return;
            (void)hipStreamSynchronize(s);
            mg_mark_idle(*res->mg);
            std::lock_guard<std::mutex> lk(sv->mu);
            if (ok) {
                res->busy = false;
                return;
            }
            for (auto it = sv->cache.begin(); it != sv->cache.end(); ++it)
                if (&*it == res) {
                    sv->reserved -= it->bytes;
                    sv->cache.erase(it);
                    break;
                }
        }
    };

    void run(Request &rq, hipStream_t s, Done &d, uint64_t &keep_bytes, uint64_t tag) {
        // resident hierarchy for this request?
        rq.key = hash_request(rq.xyz.data(), rq.xyz.size() / 3, rq.md, rq.mnp);
        Resident *res = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto it = cache.begin(); it != cache.end(); ++it)
                if (it->key == rq.key && !it->busy) {
                    it->busy = true;
                    cache.splice(cache.begin(), cache, it);
                    res = &cache.front();
                    break;
                }
        }
        ResidentLease lease;
        lease.sv = this;
        lease.res = res;
        lease.s = s;
        std::unique_ptr<MgHierarchy> own;
        MgHierarchy *mg = nullptr;

        uint64_t octants = 0;
        if (res) {
            mg = res->mg.get();
            octants = res->octants;
            d.r.reused = 1;
            bool inject = false;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (fail_next_resolve) {
                    fail_next_resolve = 0;
                    inject = true;
                }
            }
            if (inject) {
                dgx_set_error("server: injected failure of a resident re-solve");
                throw DgxError();
            }
        } else {
            const double ta = now();
            const size_t n = rq.xyz.size() / 3;
            DevBuf<double> dxyz(3 * n, s);
            if (n) HIP_OK(hipMemcpyAsync(dxyz.p, rq.xyz.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
            DevBuf<Key> u = p2o_build(dxyz.p, n, rq.md, rq.mnp, DGX_CURVE_HILBERT, s, nullptr);
            dxyz.release();
            DevBuf<Key> b = balance_build(u.p, u.n, rq.md, DGX_CURVE_HILBERT, s, nullptr);
            u.release();
            octants = b.n;
            HIP_OK(hipStreamSynchronize(s));
            const double tb = now();
            d.r.s_octree = tb - ta;
            // several slots already overlap their set-ups; one slot overlaps
            // the levels of its own set-up instead.  A request solves once, so
            // its DAs are not coloured (the colouring costs more than the
            // matvecs of one solve save).
            own = mg_build(b.p, b.n, rq.md, DGX_CURVE_HILBERT, o.nlevels, 1.0, 1.0, s, o.slots == 1, false,
                           kDenseNodesOneSolve);
            mg = own.get();
            b.release();
            HIP_OK(hipStreamSynchronize(s));
            d.r.s_hierarchy = now() - tb;
            sort_fault_check();
        }
        std::vector<double>().swap(rq.xyz);
        const double tc = now();
        const uint32_t nn = mg->lev[0].da->n_nodes;
        DevBuf<double> f(nn, s), bvec(nn, s), x(nn, s);
        if (nn) {
            k_gen_normal<<<grid_for(nn, SB), SB, 0, s>>>(rq.seed, nn, f.p);
            HIP_OK(hipGetLastError());
        }
        mg_rhs_from_nodal(*mg, f.p, bvec.p, s);
        SolveOpts so;
        so.fp32_vcycle = o.fp32_vcycle;
        so.rtol = o.rtol;
        so.maxit = o.maxit;
        // A new hierarchy is solved once: capturing and instantiating its CG
        // graphs costs more than their replay saves over a few iterations.
        // A resident one is solved again, and its graphs are captured then.
        so.use_graph = res ? 1 : 0;
        SolveStats st;
        mg_solve(*mg, bvec.p, x.p, so, &st, s);  // allocates the hierarchy's work vectors once
        d.r.iterations = st.iterations;
        d.r.rel_residual = st.rel_residual;
        d.r.octants = octants;
        d.r.nodes = nn;
        d.r.levels = (int)mg->lev.size();
        if (o.return_solution && nn) {
            d.x.resize(nn);
            HIP_OK(hipMemcpyAsync(d.x.data(), x.p, (size_t)nn * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        HIP_OK(hipStreamSynchronize(s));
        d.r.s_solve = now() - tc;
        mg_mark_idle(*mg);
        d.r.peak_bytes = (uint64_t)std::max<int64_t>(0, dgx_alloc_tag_peak(tag));
        f.release();
        bvec.release();
        x.release();
        // what is still live under this request's tag is the new hierarchy
        const int64_t hier_bytes = res ? 0 : dgx_alloc_tag_live(tag);
        if (res) {
            lease.ok = true;  // the lease clears busy once this returns
            return;
        }
        std::lock_guard<std::mutex> lk(mu);
        if (o.keep_resident) {
            Resident r;
            r.key = rq.key;
            r.bytes = (uint64_t)std::max<int64_t>(hier_bytes, 1 << 20);  // measured footprint of the hierarchy
            r.octants = octants;
            r.mg = std::move(own);
            keep_bytes = r.bytes;
            cache.push_front(std::move(r));
        }
    }
};

extern "C" {

int dgx_server_create(const dgx_server_opts *o, dgx_server *out) {
    if (!o || !out) {
        dgx_set_error("server: null argument");
        return -1;
    }
    int dev = 0;
    if (!dgx_ctx_device(&dev)) {
        dgx_set_error("dgx_init() has not been called");
        return -1;
    }
    if (o->slots < 1 || o->slots > 16 || o->nlevels < 1 || o->maxit < 1 || !(o->rtol > 0)) {
        dgx_set_error("server: bad options");
        return -1;
    }
    size_t dfree = 0, dtotal = 0;
    if (hipMemGetInfo(&dfree, &dtotal) != hipSuccess) {
        dgx_set_error("server: hipMemGetInfo failed");
        return -1;
    }
    auto *sv = new dgx_server_s();
    sv->o = *o;
    // leave a quarter of the device to the runtime, libraries and caches
    sv->o.hbm_budget = std::min<uint64_t>(o->hbm_budget, dtotal / 4 * 3);
    sv->device = dev;
    sv->t0 = std::chrono::steady_clock::now();
    for (unsigned i = 0; i < o->slots; ++i) sv->workers.emplace_back(&dgx_server_s::worker, sv, i);
    *out = sv;
    return 0;
}

int dgx_server_submit(dgx_server sv, const double *xyz, size_t n, unsigned max_depth, unsigned max_num_pts,
                      uint64_t rhs_seed, uint64_t *id) {
    if (!sv || !id || (n && !xyz)) {
        dgx_set_error("server: null argument");
        return -1;
    }
    if (max_depth > 30 || max_num_pts == 0) {
        dgx_set_error("server: maxDepth must be <= 30 and maxNumPts >= 1");
        return -1;
    }
    Request rq;
    rq.xyz.assign(xyz, xyz + 3 * n);
    rq.npts = n;
    rq.md = max_depth;
    rq.mnp = max_num_pts;
    rq.seed = rhs_seed;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        rq.est = sv->estimate(n, max_num_pts);
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        rq.id = sv->next_id++;
        rq.t_submit = sv->now();
        *id = rq.id;
        sv->queue.push_back(std::move(rq));
    }
    sv->cv_work.notify_one();
    return 0;
}

int dgx_server_wait(dgx_server sv, uint64_t id, dgx_served *res, double *x, size_t x_len) {
    if (!sv || !res) {
        dgx_set_error("server: null argument");
        return -1;
    }
    Done d;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        if (id == 0 || id >= sv->next_id) {
            dgx_set_error("server: unknown request %llu", (unsigned long long)id);
            return -1;
        }
        if (sv->consumed.count(id)) {
            dgx_set_error("server: request %llu was already waited for", (unsigned long long)id);
            return -1;
        }
        sv->cv_done.wait(lk, [&] { return sv->done.count(id) > 0 || sv->consumed.count(id) > 0; });
        if (sv->consumed.count(id)) {  // another thread's wait took it meanwhile
            dgx_set_error("server: request %llu was already waited for", (unsigned long long)id);
            return -1;
        }
        d = std::move(sv->done[id]);
        sv->done.erase(id);
        sv->consumed.insert(id);
    }
    *res = d.r;
    if (d.r.status) {
        dgx_set_error("request %llu: %s", (unsigned long long)id, d.err.c_str());
        return -1;
    }
    if (x) {
        if (d.x.size() != d.r.nodes || x_len < d.x.size()) {
            dgx_set_error("server: solution not kept (return_solution = 0) or buffer too small");
            return -1;
        }
        std::memcpy(x, d.x.data(), d.x.size() * sizeof(double));
    }
    return 0;
}

int dgx_server_peek(dgx_server sv, uint64_t id, dgx_served *res) {
    if (!sv || !res) {
        dgx_set_error("server: null argument");
        return -1;
    }
    std::unique_lock<std::mutex> lk(sv->mu);
    if (id == 0 || id >= sv->next_id) {
        dgx_set_error("server: unknown request %llu", (unsigned long long)id);
        return -1;
    }
    sv->cv_done.wait(lk, [&] { return sv->done.count(id) > 0 || sv->consumed.count(id) > 0; });
    if (sv->consumed.count(id)) {
        dgx_set_error("server: request %llu was already waited for", (unsigned long long)id);
        return -1;
    }
    *res = sv->done[id].r;
    return 0;
}

int dgx_server_get_stats(dgx_server sv, dgx_server_stats *st) {
    if (!sv || !st) {
        dgx_set_error("server: null argument");
        return -1;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    st->budget = sv->o.hbm_budget;
    st->reserved = sv->reserved;
    st->resident = sv->cache.size();
    uint64_t rb = 0;
    for (const Resident &r : sv->cache) rb += r.bytes;
    st->resident_bytes = rb;
    st->queued = sv->queue.size();
    st->running = sv->running;
    st->evictions = sv->evictions;
    return 0;
}

int dgx_server_inject_fault(dgx_server sv, int kind) {
    if (!sv || kind != DGX_SERVER_FAULT_RESOLVE) {
        dgx_set_error("server: null argument or unknown fault kind");
        return -1;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->fail_next_resolve = 1;
    return 0;
}

double dgx_server_clock(dgx_server sv) { return sv ? sv->now() : 0.0; }

int dgx_server_free(dgx_server sv) {
    if (!sv) return 0;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    for (auto &t : sv->workers) t.join();
    (void)hipSetDevice(sv->device);
    sv->cache.clear();  // frees the resident problems into the workers' streams
    sv->evicted.clear();
    for (hipStream_t st : sv->streams) {
        dgx_alloc_release_stream(st);
        (void)hipStreamDestroy(st);
    }
    delete sv;
    return 0;
}

}  // extern "C"
