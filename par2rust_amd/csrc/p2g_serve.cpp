// p2g_serve.cpp -- serving: a request loop that admits create jobs by HBM
// footprint, keeps their source and outputs resident, and runs them in
// continuous batching steps.
//
// A request brings its files (host bytes to upload, or a seed per file for
// the synthetic xorshift64 stream written on the device) and its slice size,
// recovery count and set id.  Requests are admitted in arrival order while
// their footprint (files, recovery blocks, digests) fits the budget; one that
// does not fit lets later ones that fit go ahead of it for a bounded time
// (Server::admit).  They take their memory from one pool allocated at open.
//
// A step is one batched plan (build_plan) over every admitted job with work
// left: whole jobs when their estimated device time fits the step target
// (35 ms unless the caller sets one), otherwise a piece of recovery rows
// sized to the target, or, when one row block (16 rows) is longer than what
// the whole jobs leave of the step, a run of that row block's column groups
// (the packet hashes of its rows carry their MD5 state from piece to
// piece); a job's slice checksums likewise go in as many rounds of the
// slice lanes as fit the target.  Steps run in order on one
// stream, at most two queued at a time (the next one is planned while the
// current one runs), so requests that arrive while the device is busy join
// the next step and a long job never holds the others back for more than a
// step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "p2g_batch.hpp"

namespace p2g {

constexpr uint32_t kServeSliceLanes = 24 * kBatchSliceLanes;  // slice-hash lanes of a batched launch
constexpr uint32_t kServeMinCols = 32;  // column groups (16 KiB) of the smallest column piece
// A fragments of a step: one buffer for both steps' plans, allocated at open
// (the first allocation of this size after other large frees has taken ~2 s,
// which inside the loop held every request that arrived meanwhile); a step
// whose fragments exceed it runs in more than one launch group.
constexpr uint64_t kServeAfragBytes = 16ull << 30;

// Estimated device time of one recovery row of a job and of one slice's
// hashes, from the context's measured rates (CostRates): a row block (16
// rows) of the job is S / 512 units of mfma_steps(n) K steps each, spread
// over the encode workgroups (all CUs but the hash roles' ~24); a slice is
// its MD5 blocks on one lane.
static double ms_per_row(const Ctx *c, uint32_t n, uint64_t S) {
    const double e = c->num_cus > 32 ? c->num_cus - 24.0 : c->num_cus;
    return (double)((S + 511) / 512) * mfma_steps(n) * c->rates.step_us / 16.0 / e / 1000.0;
}
static double ms_per_slice(const Ctx *c, uint64_t S) {
    return (double)((S + 8) / 64 + 1) * c->rates.slice_block_us / 1000.0;
}

// First-fit allocator over one device allocation, 256-byte granules.
struct Pool {
    uint8_t *base = nullptr;
    uint64_t size = 0, used = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    bool alloc(uint64_t n, uint64_t *off) {
        n = (n + 255) / 256 * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n, at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += n;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = (n + 255) / 256 * 256;
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    // kFailed: a step holding some of its work could not complete; its
    // outputs are not valid and p2g_server_result reports the step's error.
    enum State { kQueued, kActive, kDone, kFailed };
    uint64_t id = 0;
    State state = kQueued;
    int err = P2G_OK;  // the failed step's status
    uint64_t S = 0;
    uint32_t R = 0, n = 0;
    std::vector<uint64_t> lens, seeds;
    std::vector<const uint8_t *> host;
    uint8_t set_id[16] = {};
    uint64_t footprint = 0, off = 0;
    std::vector<p2g_dev_file> files;
    p2g_batch_job job{};
    uint32_t rows_assigned = 0, rows_done = 0;
    uint32_t slices_assigned = 0, slices_done = 0;
    // A row block too long for one step goes in as pieces of its column
    // groups over several steps: cols_assigned of the rows from
    // rows_assigned are assigned; md5_state holds their packet hashes'
    // chaining states between the pieces.
    uint32_t cols_assigned = 0;
    uint8_t *md5_state = nullptr;
    double t_submit = 0, t_admit = 0, t_done = 0;  // ms on the server clock
    double t_blocked = -1;  // when admission first found no memory for it (-1: not yet)
};

struct Step {
    BatchPlan plan;
    std::vector<std::pair<Request *, PlanItem>> items;
    bool busy = false;
    double est_ms = 0, t_launch = 0, host_ms = 0;  // P2G_SERVE_VERBOSE diagnostics
};

struct Server {
    Ctx *c = nullptr;
    Pool pool;
    double step_ms = 35.0;
    // How long a request that finds no memory may be overtaken by later
    // requests that fit (0: strict arrival order).
    double bypass_ms = 200.0;
    uint64_t next_id = 1;
    std::deque<Request *> queued;
    std::vector<Request *> active;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    std::deque<uint64_t> finished;  // done, not yet reported by poll
    Step steps[2];
    std::deque<int> inflight;  // step indices in launch order
    DevBuf afrag;              // shared by both steps' plans (one stream order)
    hipStream_t stream = nullptr;
    bool verbose = getenv("P2G_SERVE_VERBOSE") != nullptr;  // per-step timing to stderr
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    int admit();
    int retire(int k);
    int launch_step(int k);
    // Waits for every step still running that holds work of q, so its
    // memory can go back to the pool.
    int drain(const Request *q);
    ~Server();
};

Server::~Server() {
    if (c) hipSetDevice(c->device);
    if (stream) hipStreamSynchronize(stream);
    for (Step &s : steps) s.plan.afrag = &s.plan.own_afrag;
    afrag.release();
    if (pool.base) hipFree(pool.base);
    if (stream) hipStreamDestroy(stream);
}

static uint64_t a256(uint64_t x) { return (x + 255) / 256 * 256; }

// Admits queued requests in arrival order while they fit.  One that does not
// fit lets later requests that fit go ahead of it for bypass_ms from when it
// first found no memory; after that it holds back every later request until
// it is admitted, so a large request waits at most that long plus the time
// the memory it needs takes to come free.
int Server::admit() {
    const double now = now_ms();
    // the seeded files of the requests admitted now, filled by as few
    // launches as hold them (one small launch per file had spent 7% of a
    // serving run's device time in serial jump lookups)
    XsFillFiles fills{};
    auto flush = [&]() -> int {
        if (fills.n == 0) return P2G_OK;
        const hipError_t e = launch_xorshift_fill_files(
            fills, static_cast<const uint64_t *>(c->d_xs_jumps.ptr), stream);
        fills.n = 0;
        return e == hipSuccess ? P2G_OK : c->fail(P2G_ERR_DEVICE, "xorshift fill kernel", e);
    };
    for (auto it = queued.begin(); it != queued.end();) {
        Request *q = *it;
        uint64_t off = 0;
        if (!pool.alloc(q->footprint, &off)) {
            if (q->t_blocked < 0) q->t_blocked = now;
            if (now - q->t_blocked >= bypass_ms) break;
            ++it;
            continue;
        }
        it = queued.erase(it);
        q->off = off;
        uint8_t *p = pool.base + off;
        q->files.resize(q->lens.size());
        for (size_t f = 0; f < q->lens.size(); f++) {
            q->files[f] = {p, q->lens[f]};
            if (!q->host.empty() && q->host[f]) {
                const hipError_t e =
                    hipMemcpyAsync(p, q->host[f], q->lens[f], hipMemcpyHostToDevice, stream);
                if (e != hipSuccess) {
                    flush();
                    return c->fail(P2G_ERR_DEVICE, "request source upload", e);
                }
            } else if (q->lens[f]) {
                if (fills.n == kXsFillFiles) {
                    const int rc = flush();
                    if (rc != P2G_OK) return rc;
                }
                const uint32_t k = fills.n++;
                const uint64_t lanes = ((q->lens[f] + 7) / 8 + kXorshiftLaneWords - 1) / kXorshiftLaneWords;
                fills.dst[k] = p;
                fills.nbytes[k] = q->lens[f];
                fills.seed[k] = q->seeds.empty() ? 0 : q->seeds[f];
                fills.lane0[k + 1] = fills.lane0[k] + lanes;
            }
            p += a256(q->lens[f]);
        }
        q->job.files = q->files.data();
        q->job.nfiles = (uint32_t)q->files.size();
        q->job.slice_size = q->S;
        q->job.recovery_count = q->R;
        memcpy(q->job.set_id, q->set_id, 16);
        q->job.rec = p;
        p += a256((uint64_t)q->R * q->S);
        q->job.slice_md5 = p;
        p += a256(16ull * q->n);
        q->job.slice_crc = reinterpret_cast<uint32_t *>(p);
        p += a256(4ull * q->n);
        q->job.pkt_md5 = p;
        p += a256(16ull * q->R);
        q->md5_state = p;
        q->state = Request::kActive;
        q->t_admit = now_ms();
        active.push_back(q);
    }
    return flush();
}

int Server::launch_step(int k) {
    Step &st = steps[k];
    st.items.clear();
    std::vector<PlanItem> items;
    // The jobs that go whole take their share of the step first; a row block
    // longer than the step goes in as column pieces sized to what they leave
    // (at least a quarter of the target), so such a job does not stretch
    // every step past the target.
    double whole_ms = 0;
    for (Request *q : active) {
        const uint32_t left = q->R - q->rows_assigned;
        const double w = ms_per_row(c, q->n, q->S) * left;
        if (q->cols_assigned == 0 && w <= step_ms) whole_ms += w;
    }
    const double split_ms = std::max(step_ms - whole_ms, step_ms / 4);
    for (Request *q : active) {
        if (q->rows_assigned == q->R && q->slices_assigned == q->n) continue;
        const uint32_t left = q->R - q->rows_assigned;
        const uint32_t cg_total = (uint32_t)((q->S + 511) / 512);
        uint32_t take = left, c0 = 0, c1 = cg_total;
        const double row_ms = ms_per_row(c, q->n, q->S);
        if (left && (q->cols_assigned > 0 || row_ms * left > step_ms)) {
            // a piece of whole row blocks sized to the step target
            const uint32_t rows = (uint32_t)(step_ms / row_ms);
            take = std::min(left, std::max<uint32_t>(16, rows / 16 * 16));
            const uint32_t rb = std::min<uint32_t>(16, left);
            const double rb_ms = row_ms * rb;
            if (q->cols_assigned > 0 || (rb_ms > split_ms && cg_total >= 2 * kServeMinCols)) {
                // one row block is longer than the step: the next run of
                // its column groups (never a tail shorter than the minimum)
                take = rb;
                c0 = q->cols_assigned;
                const uint32_t want = std::max<uint32_t>(kServeMinCols, (uint32_t)(split_ms / rb_ms * cg_total));
                c1 = cg_total - c0 < want + kServeMinCols ? cg_total : c0 + want;
            }
        }
        // slice checksums: as many rounds of the slice lanes as fit the target
        const uint32_t sleft = q->n - q->slices_assigned;
        const uint32_t rounds = std::max<uint32_t>(1, (uint32_t)(step_ms / ms_per_slice(c, q->S)));
        const uint32_t stake = (uint32_t)std::min<uint64_t>(sleft, (uint64_t)rounds * kServeSliceLanes);
        PlanItem it{&q->job, q->rows_assigned, q->rows_assigned + take, q->slices_assigned,
                    q->slices_assigned + stake};
        it.c0 = c0;
        it.c1 = c1;
        it.md5_state = q->md5_state;
        if (c1 == cg_total) {
            q->rows_assigned += take;
            q->cols_assigned = 0;
        } else {
            q->cols_assigned = c1;
        }
        q->slices_assigned += stake;
        st.items.push_back({q, it});
        items.push_back(it);
    }
    if (items.empty()) return P2G_OK;
    const double t_plan = now_ms();
    int rc = build_plan(c, items, false, &st.plan);
    if (rc == P2G_OK) rc = launch_plan(&st.plan, stream);
    st.t_launch = now_ms();
    st.host_ms = st.t_launch - t_plan;
    st.est_ms = 0;
    for (auto &pr : st.items) {
        const Request *q = pr.first;
        const double cols = (double)((q->S + 511) / 512);
        st.est_ms += ms_per_row(c, q->n, q->S) * (pr.second.r1 - pr.second.r0) *
                     (std::min<double>(pr.second.c1, cols) - pr.second.c0) / cols;
    }
    if (rc != P2G_OK) {
        // whatever part of the step was queued ends before its requests are
        // reported failed (their memory may then go back to the pool)
        hipStreamSynchronize(stream);
        const double t = now_ms();
        for (auto &pr : st.items) {
            Request *q = pr.first;
            if (q->state != Request::kActive) continue;
            q->state = Request::kFailed;
            q->err = rc;
            q->t_done = t;
            finished.push_back(q->id);
            active.erase(std::find(active.begin(), active.end(), q));
        }
        st.items.clear();
        return rc;
    }
    st.busy = true;
    inflight.push_back(k);
    return P2G_OK;
}

int Server::retire(int k) {
    Step &st = steps[k];
    const int rc = wait_plan(&st.plan);
    st.busy = false;
    const double t = now_ms();
    if (verbose && rc == P2G_OK && st.plan.h_done.size() == st.plan.njobs + 1) {
        // device start (first batched workgroup) and end (last completion
        // stamp) on the 100 MHz counter, host planning time, estimate
        const unsigned long long t0 = st.plan.h_done[st.plan.njobs];
        unsigned long long t1 = t0;
        for (uint32_t j = 0; j < st.plan.njobs; j++) t1 = std::max(t1, st.plan.h_done[j]);
        fprintf(stderr, "serve step: items %zu groups %zu est %.1f ms | device %.3f-%.3f ms (%.1f) | host plan %.1f ms at %.1f, retired at %.1f\n",
                st.items.size(), st.plan.groups.size(), st.est_ms, t0 * 1e-5, t1 * 1e-5,
                (t1 - t0) * 1e-5, st.host_ms, st.t_launch, t);
    }
    for (auto &pr : st.items) {
        Request *q = pr.first;
        if (q->state != Request::kActive) continue;  // failed in an earlier step
        if (rc != P2G_OK) {
            // the step's work did not complete (a packet hash gave up, or the
            // launch failed): none of it counts, and the request is reported
            // as failed rather than done
            q->state = Request::kFailed;
            q->err = rc;
            q->t_done = t;
            finished.push_back(q->id);
            active.erase(std::find(active.begin(), active.end(), q));
            continue;
        }
        // rows count once their last column piece is done
        if (pr.second.c1 >= (q->S + 511) / 512) q->rows_done += pr.second.r1 - pr.second.r0;
        q->slices_done += pr.second.s1 - pr.second.s0;
        if (q->rows_done == q->R && q->slices_done == q->n) {
            q->state = Request::kDone;
            q->t_done = t;
            finished.push_back(q->id);
            active.erase(std::find(active.begin(), active.end(), q));
        }
    }
    st.items.clear();
    // buffers the plans outgrew go back while no step is queued (a free
    // waits for the whole device)
    if (!steps[0].busy && !steps[1].busy)
        for (Step &s : steps)
            if (!s.plan.retired_dev.empty() || !s.plan.retired_host.empty()) s.plan.free_retired();
    return rc;
}

int Server::drain(const Request *q) {
    for (Step &st : steps) {
        if (!st.busy) continue;
        for (auto &pr : st.items)
            if (pr.first == q) {
                const hipError_t e = hipEventSynchronize(st.plan.done);
                if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "wait for a step", e);
                break;
            }
    }
    return P2G_OK;
}

}  // namespace p2g

using namespace p2g;

extern "C" {

int p2g_server_open(p2g_ctx *ctx, uint64_t hbm_budget, double step_ms, p2g_server **out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c || !out || hbm_budget == 0) return P2G_ERR_INVALID_ARGUMENT;
        *out = nullptr;
        hipSetDevice(c->device);
        Server *s = new (std::nothrow) Server();
        if (!s) return c->fail(P2G_ERR_OUT_OF_MEMORY, "server");
        s->c = c;
        if (step_ms > 0) s->step_ms = step_ms;
        s->bypass_ms = 4 * s->step_ms;
        // the fill tables of seeded requests, built before any step is queued
        if (const int rc = xs_jump_tables(c); rc != P2G_OK) {
            delete s;
            return rc;
        }
        s->pool.size = a256(hbm_budget);
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, s->pool.size);
        if (e != hipSuccess) {
            delete s;
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "server HBM pool", e);
        }
        s->pool.base = static_cast<uint8_t *>(p);
        s->pool.free_[0] = s->pool.size;
        e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete s;
            return c->fail(P2G_ERR_DEVICE, "server stream", e);
        }
        // the steps' A fragments, smaller when the device is short of memory
        uint64_t afrag_bytes = kServeAfragBytes;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            afrag_bytes = std::max<uint64_t>(256ull << 20, std::min<uint64_t>(afrag_bytes, free_b / 2));
        if (s->afrag.reserve(afrag_bytes) != P2G_OK) {
            delete s;
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "server fragment buffer");
        }
        for (Step &st : s->steps) {
            st.plan.afrag = &s->afrag;
            st.plan.afrag_budget = afrag_bytes;
            st.plan.keep_retired = true;
        }
        *out = reinterpret_cast<p2g_server *>(s);
        return P2G_OK;
    });
}

int p2g_server_set_bypass(p2g_server *srv, double bypass_ms) {
    return guarded(srv ? reinterpret_cast<Server *>(srv)->c : nullptr, [&]() -> int {
        Server *s = reinterpret_cast<Server *>(srv);
        if (!s || !(bypass_ms >= 0)) return P2G_ERR_INVALID_ARGUMENT;
        s->bypass_ms = bypass_ms;
        return P2G_OK;
    });
}

int p2g_server_submit(p2g_server *srv, const p2g_request *r, uint64_t *id_out) {
    return guarded(srv ? reinterpret_cast<Server *>(srv)->c : nullptr, [&]() -> int {
        Server *s = reinterpret_cast<Server *>(srv);
        if (!s || !r || !id_out) return P2G_ERR_INVALID_ARGUMENT;
        Ctx *c = s->c;
        if (r->nfiles && !r->file_lens) return c->fail(P2G_ERR_INVALID_ARGUMENT, "null file lengths");
        int rc = check_slice_size(c, r->slice_size);
        if (rc != P2G_OK) return rc;
        if (r->recovery_count > P2G_MAX_RECOVERY_BLOCKS)
            return c->fail(P2G_ERR_TOO_MANY_RECOVERY_BLOCKS, "recovery count above 65535");
        auto q = std::make_unique<Request>();
        q->S = r->slice_size;
        q->R = r->recovery_count;
        uint64_t n = 0, bytes = 0;
        for (uint32_t f = 0; f < r->nfiles; f++) {
            q->lens.push_back(r->file_lens[f]);
            n += (r->file_lens[f] + q->S - 1) / q->S;
            bytes += a256(r->file_lens[f]);
            if (r->host_data) q->host.push_back(r->host_data[f]);
            q->seeds.push_back(r->fill_seeds ? r->fill_seeds[f] : 0);
        }
        if (n == 0) return c->fail(P2G_ERR_INVALID_ARGUMENT, "a request has no slices");
        if (n > kMaxInputs) return c->fail(P2G_ERR_TOO_MANY_INPUT_BLOCKS, "more than 32768 input slices");
        q->n = (uint32_t)n;
        memcpy(q->set_id, r->set_id, 16);
        q->footprint = bytes + a256((uint64_t)q->R * q->S) + a256(16 * n) + a256(4 * n) +
                       2 * a256(16ull * q->R);  // packet digests, then their chaining states
        if (q->footprint > s->pool.size)
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "request larger than the server's HBM budget");
        q->id = s->next_id++;
        q->t_submit = s->now_ms();
        *id_out = q->id;
        s->queued.push_back(q.get());
        s->all[q->id] = std::move(q);
        return P2G_OK;
    });
}

int p2g_server_poll(p2g_server *srv, double wait_ms, uint64_t *done_ids, uint32_t cap,
                    uint32_t *n_done) {
    return guarded(srv ? reinterpret_cast<Server *>(srv)->c : nullptr, [&]() -> int {
        Server *s = reinterpret_cast<Server *>(srv);
        if (!s || (cap && !done_ids)) return P2G_ERR_INVALID_ARGUMENT;
        hipSetDevice(s->c->device);
        const double until = s->now_ms() + std::max(0.0, wait_ms);
        int rc = P2G_OK;
        for (;;) {
            // retire finished steps, in launch order
            while (!s->inflight.empty()) {
                const int k = s->inflight.front();
                const hipError_t q = hipEventQuery(s->steps[k].plan.done);
                if (q == hipErrorNotReady) break;
                s->inflight.pop_front();
                const int r = s->retire(k);
                if (rc == P2G_OK) rc = r;
            }
            if (rc == P2G_OK) rc = s->admit();
            // keep one step running and the next one queued behind it
            while (rc == P2G_OK && s->inflight.size() < 2) {
                int k = !s->steps[0].busy ? 0 : !s->steps[1].busy ? 1 : -1;
                if (k < 0) break;
                const size_t before = s->inflight.size();
                rc = s->launch_step(k);
                if (s->inflight.size() == before) break;  // nothing to run
            }
            if (rc != P2G_OK || !s->finished.empty() || s->now_ms() >= until) break;
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        // On an error the ids stay queued for the next poll (the requests of a
        // failed step are among them, reported as failed by result).
        uint32_t k = 0;
        while (rc == P2G_OK && k < cap && !s->finished.empty()) {
            done_ids[k++] = s->finished.front();
            s->finished.pop_front();
        }
        if (n_done) *n_done = k;
        return rc;
    });
}

int p2g_server_result(p2g_server *srv, uint64_t id, p2g_batch_job *out, double times_ms[3]) {
    return guarded(srv ? reinterpret_cast<Server *>(srv)->c : nullptr, [&]() -> int {
        Server *s = reinterpret_cast<Server *>(srv);
        if (!s) return P2G_ERR_INVALID_ARGUMENT;
        auto it = s->all.find(id);
        if (it == s->all.end()) return s->c->fail(P2G_ERR_INVALID_ARGUMENT, "unknown request id");
        Request *q = it->second.get();
        if (q->state != Request::kDone && q->state != Request::kFailed)
            return s->c->fail(P2G_ERR_INVALID_ARGUMENT, "request not done");
        if (out) *out = q->job;
        if (times_ms) {
            times_ms[0] = q->t_submit;
            times_ms[1] = q->t_admit;
            times_ms[2] = q->t_done;
        }
        if (q->state == Request::kFailed)
            return s->c->fail(q->err, "request failed: a step holding its work could not complete");
        return P2G_OK;
    });
}

int p2g_server_release(p2g_server *srv, uint64_t id) {
    return guarded(srv ? reinterpret_cast<Server *>(srv)->c : nullptr, [&]() -> int {
        Server *s = reinterpret_cast<Server *>(srv);
        if (!s) return P2G_ERR_INVALID_ARGUMENT;
        auto it = s->all.find(id);
        if (it == s->all.end()) return s->c->fail(P2G_ERR_INVALID_ARGUMENT, "unknown request id");
        Request *q = it->second.get();
        if (q->state != Request::kDone && q->state != Request::kFailed)
            return s->c->fail(P2G_ERR_INVALID_ARGUMENT, "request not done");
        // a failed request may still have pieces in the step queued after
        // the failed one: its memory returns to the pool once that ends
        hipSetDevice(s->c->device);
        const int rc = s->drain(q);
        if (rc != P2G_OK) return rc;
        s->pool.release(q->off, q->footprint);
        s->all.erase(it);
        return P2G_OK;
    });
}

int p2g_server_stats(p2g_server *srv, uint32_t *queued, uint32_t *active, uint64_t *hbm_used) {
    return guarded(srv ? reinterpret_cast<Server *>(srv)->c : nullptr, [&]() -> int {
        Server *s = reinterpret_cast<Server *>(srv);
        if (!s) return P2G_ERR_INVALID_ARGUMENT;
        if (queued) *queued = (uint32_t)s->queued.size();
        if (active) *active = (uint32_t)s->active.size();
        if (hbm_used) *hbm_used = s->pool.used;
        return P2G_OK;
    });
}

void p2g_server_close(p2g_server *srv) { delete reinterpret_cast<Server *>(srv); }

}  // extern "C"
