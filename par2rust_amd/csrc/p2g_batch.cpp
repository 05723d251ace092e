// p2g_batch.cpp -- batched create: many independent create jobs (each the work
// of p2g_dev_create_pass) in persistent launches that share the GPU.
//
// Planning (host): every job's slice descriptors go into one table, ordered by
// job (longest slices first, so the longest serial hashes start first; then
// most inputs first).  Partial slices get zero-padded copies, so every
// descriptor is a whole slice.  Each job's recovery rows are cut into pieces
// whose A fragments fit the group budget, and pieces are packed in order
// into groups within the fragment budget (below); a group is
// one k_create_batch launch after one k_batch_afrag launch.  A job's slice
// checksums run in the group of its first piece.
//
// The launch sequence (staging copies, descriptor tails, then per group the
// fragments and the pass) is recorded once into a HIP graph and replayed by
// p2g_batch_launch; a plan made with use_graph = 0 issues the launches
// directly each time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "p2g_batch.hpp"

namespace p2g {

// A fragments of one group: a quarter of the device memory free at planning
// time, within [16, 64] GiB (P2G_BATCH_AFRAG_BYTES overrides it); one row
// block of a job is the smallest piece, whatever the budget.  Fewer groups
// mean fewer fragment launches and group tails: the 256-job mix ran in 2
// groups at 64 GiB, 0.399 against 0.402-0.404 s per replay in 6 at 16 GiB.
constexpr uint64_t kBatchAfragMin = 16ull << 30, kBatchAfragMax = 64ull << 30;

void BatchPlan::drop_graph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
}

void BatchPlan::free_retired() {
    for (void *p : retired_dev) hipFree(p);
    for (void *p : retired_host) hipHostFree(p);
    retired_dev.clear();
    retired_host.clear();
    retired_bytes = 0;
}

BatchPlan::~BatchPlan() {
    if (c) hipSetDevice(c->device);
    if (done) hipEventSynchronize(done);
    drop_graph();
    if (cap) hipStreamDestroy(cap);
    if (done) hipEventDestroy(done);
    for (DevBuf *b : {&d_descs, &d_tb, &d_staging, &d_parts, &d_pieces, &d_sjobs, &d_sync, &d_done,
                      &d_err, &own_afrag})
        b->release();
    free_retired();
}

int BatchPlan::record(hipStream_t st) {
    hipError_t e = launch_stage_parts(static_cast<const StagePart *>(d_parts.ptr), n_parts,
                                      max_part_S, st);
    if (e == hipSuccess)
        e = launch_desc_tails(static_cast<SliceDesc *>(d_descs.ptr), n_slices,
                              static_cast<uint4 *>(d_tb.ptr), st);
    for (const BatchGroupLaunch &g : groups) {
        if (e == hipSuccess)
            e = launch_batch_afrag(static_cast<const uint16_t *>(c->d_antilog.ptr),
                                   static_cast<const uint16_t *>(c->d_base_logs.ptr), g, st);
        if (e == hipSuccess)
            e = launch_create_batch(g, static_cast<const SliceDesc *>(d_descs.ptr),
                                    static_cast<const uint4 *>(d_tb.ptr), st);
    }
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "batched create launches", e);
    return P2G_OK;
}

void free_batch_plan(BatchPlan *b) { delete b; }

static void packet_prefix_words(const uint8_t set_id[16], uint32_t out[8]) {
    static const uint8_t type_tag[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                         'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
    for (int w = 0; w < 4; w++) {
        memcpy(&out[w], set_id + 4 * w, 4);  // little-endian host, as the device
        memcpy(&out[4 + w], type_tag + 4 * w, 4);
    }
}

static uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// Buffers of a plan grow with room to spare (half again, at most 1 GiB more),
// so a plan rebuilt for steps of varying size (the serving loop) rarely
// reallocates.  Freeing the old buffer waits for the whole device, unless
// the plan keeps it for later (keep_retired), up to kRetiredMax bytes kept.
constexpr uint64_t kRetiredMax = 4ull << 30;

// Whether the old buffer (cap bytes) may wait in the plan; if not, what the
// plan kept is freed now, after the device is idle.
static bool keep_old(BatchPlan *B, size_t cap) {
    if (!B->keep_retired) return false;
    if (B->retired_bytes + cap <= kRetiredMax) return true;
    hipDeviceSynchronize();
    B->free_retired();
    return false;
}
static int grow(BatchPlan *B, DevBuf &d, size_t n) {
    if (n <= d.cap) return P2G_OK;
    const size_t want = n + std::min<size_t>(n / 2, 1ull << 30);
    if (!d.ptr || !keep_old(B, d.cap)) return d.reserve(want);
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) return P2G_ERR_OUT_OF_MEMORY;
    B->retired_dev.push_back(d.ptr);
    B->retired_bytes += d.cap;
    d.ptr = p;
    d.cap = want;
    return P2G_OK;
}
static int grow(BatchPlan *B, PinBuf &h, size_t n) {
    if (n <= h.cap) return P2G_OK;
    const size_t want = n + std::min<size_t>(n / 2, 1ull << 30);
    if (!h.ptr || !keep_old(B, h.cap)) return h.reserve(want);
    void *p = nullptr;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) return P2G_ERR_OUT_OF_MEMORY;
    B->retired_host.push_back(h.ptr);
    B->retired_bytes += h.cap;
    h.ptr = p;
    h.cap = want;
    return P2G_OK;
}

// A table for the next launch to upload: the device buffer is sized now,
// the bytes go to the pinned staging area at `off` (reserved by the caller).
template <typename T>
static int stage(Ctx *c, BatchPlan *B, DevBuf &d, const std::vector<T> &v, size_t off,
                 const char *what) {
    if (grow(B, d, std::max<size_t>(v.size() * sizeof(T), 16)) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, what);
    if (v.empty()) return P2G_OK;
    memcpy(static_cast<uint8_t *>(B->h_stage.ptr) + off, v.data(), v.size() * sizeof(T));
    B->pending.push_back({&d, off, v.size() * sizeof(T)});
    return P2G_OK;
}

int check_batch_job(Ctx *c, const p2g_batch_job &J, std::vector<SliceDesc> &descs) {
    const uint32_t R = J.recovery_count;
    if ((J.nfiles && !J.files) || !J.slice_md5 || !J.slice_crc || (R && (!J.rec || !J.pkt_md5)))
        return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument in a batch job");
    for (const void *p : {(const void *)J.rec, (const void *)J.slice_md5, (const void *)J.slice_crc,
                          (const void *)J.pkt_md5})
        if (reinterpret_cast<uintptr_t>(p) % 4 != 0)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "batch outputs must be 4-byte aligned");
    if (R > P2G_MAX_RECOVERY_BLOCKS)
        return c->fail(P2G_ERR_TOO_MANY_RECOVERY_BLOCKS, "recovery count above 65535");
    int rc = check_slice_size(c, J.slice_size);
    if (rc != P2G_OK) return rc;
    descs.clear();
    rc = append_descs(c, J.files, J.nfiles, J.slice_size, descs);
    if (rc != P2G_OK) return rc;
    if (descs.empty()) return c->fail(P2G_ERR_INVALID_ARGUMENT, "a batch job has no slices");
    if (descs.size() > kMaxInputs)
        return c->fail(P2G_ERR_TOO_MANY_INPUT_BLOCKS, "more than 32768 input slices");
    return P2G_OK;
}

// The sizing work of cnt slices of S bytes (one lane's MD5 + CRC-32 each).
static void add_slice_work(BatchGroupWork &w, uint32_t cnt, uint32_t S) {
    const double blocks = (double)((S + 8) / 64 + 1);  // MD5 blocks of a slice
    w.slices += cnt;
    w.slice_blocks += cnt * blocks;
    if (cnt && blocks > w.big_blocks) {
        w.big_blocks = blocks;
        w.big_slices = 0;
    }
    if (cnt && blocks == w.big_blocks) w.big_slices += cnt;
    if (cnt && (w.small_blocks == 0 || blocks < w.small_blocks)) w.small_blocks = blocks;
}

// The sizing work of one piece: its encode unit-steps and packet tasks.
// Blocks of a packet body's hash within data bytes [0, e) (as the packet
// lanes count them).
static uint32_t packet_blocks_within(uint64_t e) { return e >= 28 ? (uint32_t)((e - 28) / 64) + 1 : 0u; }

static void add_piece_work(BatchGroupWork &w, const BatchPiece &P) {
    // the packet hash blocks the piece's column groups hold
    const uint64_t nb = (36ull + P.S + 8) / 64 + 1;
    const uint64_t e1 = std::min<uint64_t>(512ull * (P.cg0 + P.col_groups), P.S);
    const uint64_t hi = P.cg0 + P.col_groups >= P.cg_total ? nb : packet_blocks_within(e1);
    const uint64_t lo = P.cg0 ? packet_blocks_within(512ull * P.cg0) : 0;
    const double blocks = (double)(hi > lo ? hi - lo : 1);
    const uint32_t t = (P.R + 63) / 64;
    const double steps = (double)P.row_blocks * P.col_groups * P.nq;
    w.enc_steps += steps;
    w.piece_steps.push_back(steps);
    w.piece_tasks.push_back(t);
    w.piece_task_blocks.push_back(blocks);
}

// Plans the items into B (its buffers are reused and grown; the previous
// launch of B must have finished).  Completion stamps are per item.
int build_plan(Ctx *c, const std::vector<PlanItem> &items, bool use_graph, BatchPlan *B) {
    const uint32_t nitems = (uint32_t)items.size();
    if (nitems == 0) return c->fail(P2G_ERR_INVALID_ARGUMENT, "empty batch");
    hipSetDevice(c->device);
    B->c = c;
    B->njobs = nitems;
    B->groups.clear();
    B->work.clear();
    B->sync_off.clear();
    B->drop_graph();
    std::vector<std::vector<SliceDesc>> jd(nitems);
    for (uint32_t j = 0; j < nitems; j++) {
        const PlanItem &it = items[j];
        int rc = check_batch_job(c, *it.job, jd[j]);
        if (rc != P2G_OK) return rc;
        if (it.r0 > it.r1 || it.r1 > it.job->recovery_count)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "recovery row range of a batch item");
        const uint64_t cg_total = (it.job->slice_size + 511) / 512;
        if (it.c0 >= std::min<uint64_t>(it.c1, cg_total) && it.r0 < it.r1)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "column range of a batch item");
        if ((it.c0 > 0 || it.c1 < cg_total) && it.r0 < it.r1 && !it.md5_state)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "a batch item of part of a row needs its hash state");
    }
    std::vector<uint32_t> order(nitems);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const uint64_t sa = items[a].job->slice_size, sb = items[b].job->slice_size;
        if (sa != sb) return sa > sb;
        return jd[a].size() > jd[b].size();
    });

    // staging slots of the partial slices
    uint64_t stage_bytes = 0, n_slices = 0;
    for (uint32_t j : order) {
        const uint32_t S = (uint32_t)items[j].job->slice_size;
        for (const SliceDesc &d : jd[j])
            if (d.valid < S) stage_bytes = align_up(stage_bytes, 256) + S;
        n_slices += jd[j].size();
    }
    if (n_slices > 0x7FFFFFFFull) return c->fail(P2G_ERR_INVALID_ARGUMENT, "too many slices");
    if (grow(B, B->d_staging, std::max<uint64_t>(stage_bytes, 256)) != P2G_OK ||
        grow(B, B->d_descs, n_slices * sizeof(SliceDesc)) != P2G_OK ||
        grow(B, B->d_tb, n_slices * 64) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "batch descriptor tables");
    uint8_t *staging = static_cast<uint8_t *>(B->d_staging.ptr);
    const SliceDesc *d_descs = static_cast<const SliceDesc *>(B->d_descs.ptr);

    std::vector<SliceDesc> all;
    all.reserve(n_slices);
    std::vector<StagePart> parts;
    std::vector<BatchSlices> sjobs;
    std::vector<uint32_t> slice0(nitems);
    std::vector<uint8_t> has_slices(nitems);
    uint64_t so = 0;
    B->max_part_S = 0;
    for (uint32_t j : order) {
        const p2g_batch_job &J = *items[j].job;
        const uint32_t S = (uint32_t)J.slice_size;
        slice0[j] = (uint32_t)all.size();
        for (const SliceDesc &d : jd[j]) {
            if (d.valid < S) {
                so = align_up(so, 256);
                parts.push_back({d.ptr, staging + so, d.valid, S, 0});
                all.push_back({staging + so, S, 0});
                B->max_part_S = std::max(B->max_part_S, S);
                so += S;
            } else {
                all.push_back(d);
            }
        }
        // the item's slice range (possibly empty) within the job's slices
        const uint32_t n = (uint32_t)jd[j].size();
        const uint32_t s0 = std::min(items[j].s0, n), s1 = std::max(s0, std::min(items[j].s1, n));
        sjobs.push_back({J.slice_md5 + 16ull * s0, J.slice_crc + s0, slice0[j] + s0, slice0[j] + s1,
                         S, j});
    }
    B->n_slices = (uint32_t)n_slices;
    B->n_parts = (uint32_t)parts.size();

    // pieces and groups
    const uint32_t band_default = c->band_cols ? c->band_cols : 64u;
    uint64_t budget = B->afrag_budget ? B->afrag_budget : kBatchAfragMin;
    if (const char *v = getenv("P2G_BATCH_AFRAG_BYTES")) {
        budget = std::max<uint64_t>(1, strtoull(v, nullptr, 10));
    } else if (!B->afrag_budget) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            budget = std::min(kBatchAfragMax, std::max<uint64_t>(kBatchAfragMin, free_b / 4));
    }
    std::vector<BatchPiece> pieces;
    std::vector<uint64_t> afrag_off;  // of each piece inside its group's fragments
    struct GroupSpan {
        uint32_t p_lo, p_hi, s_lo, s_hi;
        uint64_t afrag;
    };
    std::vector<GroupSpan> spans(1, GroupSpan{0, 0, 0, 0, 0});
    std::vector<uint32_t> job_group(nitems);  // by position in `order`; non-decreasing
    uint64_t afrag_max = 0;
    for (uint32_t oi = 0; oi < nitems; oi++) {
        const uint32_t j = order[oi];
        const p2g_batch_job &J = *items[j].job;
        const uint32_t S = (uint32_t)J.slice_size, n = (uint32_t)jd[j].size();
        const uint32_t rlo = items[j].r0, rhi = items[j].r1, nq = mfma_steps(n);
        const uint64_t per_rb = mfma_afrag_bytes(n, 16);  // fragments of 16 recovery rows
        // pieces of equal row-block counts, each within the budget
        const uint32_t max_rb = (uint32_t)std::max<uint64_t>(1, budget / per_rb);
        const uint32_t rbs = (rhi - rlo + 15) / 16, npieces = (rbs + max_rb - 1) / max_rb;
        const uint32_t piece_rows = 16 * (npieces ? (rbs + npieces - 1) / npieces : 1);
        job_group[oi] = (uint32_t)spans.size() - 1;
        for (uint32_t r0 = rlo; r0 < rhi; r0 += piece_rows) {
            const uint32_t Rp = std::min<uint32_t>(piece_rows, rhi - r0);
            const uint64_t bytes = mfma_afrag_bytes(n, Rp);
            if (spans.back().p_hi > spans.back().p_lo && spans.back().afrag + bytes > budget) {
                const uint32_t np = (uint32_t)pieces.size();
                spans.push_back({np, np, 0, 0, 0});
            }
            if (r0 == rlo) job_group[oi] = (uint32_t)spans.size() - 1;
            BatchPiece P{};
            P.inputs = d_descs + slice0[j];
            P.rec = J.rec + (size_t)r0 * S;
            P.pkt_md5 = J.pkt_md5 + (size_t)r0 * 16;
            packet_prefix_words(J.set_id, P.pre);
            P.n = n;
            P.nq = nq;
            P.R = Rp;
            P.S = S;
            P.first_exp = r0;
            P.job = j;
            P.row_blocks = (Rp + 15) / 16;
            P.cg_total = (S + 511) / 512;
            P.cg0 = items[j].c0;
            P.col_groups = std::min(items[j].c1, P.cg_total) - P.cg0;
            P.md5_state = items[j].md5_state
                              ? static_cast<uint4 *>(items[j].md5_state) + r0
                              : nullptr;
            TuneChoice ch;
            const uint32_t band = tune_lookup(c, n, J.recovery_count, S, &ch) &&
                                          ch.variant == kEncodeMfma && ch.band_cols
                                      ? ch.band_cols
                                      : band_default;
            // rows of at most 32 column groups (slices up to 16 KiB) are one band
            P.band_cols = P.col_groups <= 32 ? P.col_groups
                                             : std::max<uint32_t>(1, std::min(band, P.col_groups));
            P.n_bands = 0;
            for (uint32_t cg = 0; cg < P.col_groups; P.n_bands++)
                cg += batch_band_width(cg, P.col_groups - cg, P.band_cols);
            pieces.push_back(P);
            afrag_off.push_back(spans.back().afrag);
            spans.back().p_hi = (uint32_t)pieces.size();
            spans.back().afrag += bytes;
            afrag_max = std::max(afrag_max, spans.back().afrag);
        }
    }
    // the slice jobs of group k: the items whose first piece (or, with no
    // recovery rows, whose place in the order) falls in it
    for (uint32_t k = 0; k < spans.size(); k++) {
        spans[k].s_lo = (uint32_t)(std::lower_bound(job_group.begin(), job_group.end(), k) - job_group.begin());
        spans[k].s_hi = (uint32_t)(std::upper_bound(job_group.begin(), job_group.end(), k) - job_group.begin());
    }

    if (grow(B, *B->afrag, std::max<uint64_t>(afrag_max, 256)) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "batch A fragments");
    uint8_t *afrag_base = static_cast<uint8_t *>(B->afrag->ptr);

    // per group: units, tasks, sync layout
    uint32_t sync_words_total = 0;
    std::vector<BatchGroupLaunch> gl(spans.size());
    for (size_t k = 0; k < spans.size(); k++) {
        const GroupSpan &sp = spans[k];
        BatchGroupLaunch &g = gl[k];
        g = BatchGroupLaunch{};
        uint32_t units = 0, tasks = 0, bands = 0;
        for (uint32_t p = sp.p_lo; p < sp.p_hi; p++) {
            BatchPiece &P = pieces[p];
            P.afrag = afrag_base + afrag_off[p];
            P.unit0 = units;
            units += P.row_blocks * P.col_groups;
            P.unit_end = units;
            P.task0 = tasks;
            tasks += (P.R + 63) / 64;
            P.task_end = tasks;
            P.band0 = kPassBands + bands;
            bands += P.n_bands;
        }
        uint32_t starts = kPassBands + bands;
        for (uint32_t p = sp.p_lo; p < sp.p_hi; p++) {
            pieces[p].bstart0 = starts;
            starts += pieces[p].n_bands + 1;
        }
        g.n_pieces = sp.p_hi - sp.p_lo;
        g.n_sjobs = sp.s_hi - sp.s_lo;
        g.slice_lo = sp.s_lo < sp.s_hi ? sjobs[sp.s_lo].slice0 : 0;
        g.slice_hi = sp.s_lo < sp.s_hi ? sjobs[sp.s_hi - 1].slice_end : 0;
        g.n_units = units;
        g.n_tasks = tasks;
        g.sync_words = kPassBands + bands;
        g.n_jobs = nitems;
        g.first = k == 0;
        g.afrag_entries = sp.afrag / 16;
        // [slice_lo, slice_hi) may hold the descriptors of items without
        // checksums; the lanes skip them, and only checksummed slices count
        uint32_t hashed = 0;
        BatchGroupWork w;
        for (uint32_t q = sp.s_lo; q < sp.s_hi; q++) {
            const uint32_t cnt = sjobs[q].slice_end - sjobs[q].slice0;
            hashed += cnt;
            add_slice_work(w, cnt, sjobs[q].S);
        }
        for (uint32_t p = sp.p_lo; p < sp.p_hi; p++) add_piece_work(w, pieces[p]);
        plan_batch_group(c->num_cus, c->batch_resident, tasks, hashed, units, w, c->rates, &g);
        B->work.push_back(w);
        B->sync_off.push_back(sync_words_total);
        sync_words_total += (uint32_t)align_up(starts, 64);
    }
    // the tables, staged in pinned memory for the next launch to upload
    const size_t o_descs = 0, o_parts = align_up(o_descs + all.size() * sizeof(SliceDesc), 256);
    const size_t o_pieces = align_up(o_parts + parts.size() * sizeof(StagePart), 256);
    const size_t o_sjobs = align_up(o_pieces + pieces.size() * sizeof(BatchPiece), 256);
    const size_t stage_total = o_sjobs + sjobs.size() * sizeof(BatchSlices);
    B->pending.clear();
    if (grow(B, B->h_stage, std::max<size_t>(stage_total, 256)) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "batch table staging");
    int rc = stage(c, B, B->d_descs, all, o_descs, "batch descriptors");
    if (rc == P2G_OK) rc = stage(c, B, B->d_parts, parts, o_parts, "batch staging table");
    if (rc == P2G_OK) rc = stage(c, B, B->d_pieces, pieces, o_pieces, "batch pieces");
    if (rc == P2G_OK) rc = stage(c, B, B->d_sjobs, sjobs, o_sjobs, "batch slice jobs");
    if (rc != P2G_OK) return rc;
    const bool new_err = B->d_err.ptr == nullptr;
    if (grow(B, B->d_sync, (size_t)sync_words_total * 4) != P2G_OK ||
        grow(B, B->d_done, (size_t)(nitems + 1) * 8) != P2G_OK || B->d_err.reserve(256) != P2G_OK ||
        grow(B, B->h_back, 8 + (size_t)(nitems + 1) * 8 + spans.size() * 8 * sizeof(unsigned long long)) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "batch sync words");
    if (new_err) {
        hipError_t ez = hipMemset(B->d_err.ptr, 0, 256);
        if (ez != hipSuccess) return c->fail(P2G_ERR_DEVICE, "clear batch error word", ez);
    }
    for (size_t k = 0; k < gl.size(); k++) {
        gl[k].pieces = static_cast<const BatchPiece *>(B->d_pieces.ptr) + spans[k].p_lo;
        gl[k].sjobs = static_cast<const BatchSlices *>(B->d_sjobs.ptr) + spans[k].s_lo;
        gl[k].sync = static_cast<uint32_t *>(B->d_sync.ptr) + B->sync_off[k];
        gl[k].err = static_cast<uint32_t *>(B->d_err.ptr);
        gl[k].job_done = static_cast<unsigned long long *>(B->d_done.ptr);
    }
    B->groups = gl;
    hipError_t e = hipSuccess;
    if (!B->done) e = hipEventCreateWithFlags(&B->done, hipEventDisableTiming);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "hipEventCreate", e);
    if (use_graph) {
        if (!B->cap) e = hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "begin graph capture", e);
        rc = B->record(B->cap);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(B->cap, &graph);
        B->graph = graph;
        if (rc != P2G_OK) return rc;
        if (e == hipSuccess) e = hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "batch graph", e);
    }
    return P2G_OK;
}

static int make_plan(Ctx *c, const p2g_batch_job *jobs, uint32_t njobs, bool use_graph,
                     BatchPlan **out) {
    *out = nullptr;
    if (njobs == 0 || !jobs) return c->fail(P2G_ERR_INVALID_ARGUMENT, "empty batch");
    std::vector<PlanItem> items(njobs);
    for (uint32_t j = 0; j < njobs; j++) items[j] = {&jobs[j], 0, jobs[j].recovery_count, 0, UINT32_MAX};
    BatchPlan *B = new (std::nothrow) BatchPlan();
    if (!B) return c->fail(P2G_ERR_OUT_OF_MEMORY, "batch plan");
    const int rc = build_plan(c, items, use_graph, B);
    if (rc != P2G_OK) {
        delete B;
        return rc;
    }
    *out = B;
    return P2G_OK;
}

int launch_plan(BatchPlan *B, hipStream_t st) {
    Ctx *c = B->c;
    hipSetDevice(c->device);
    // launches of one plan share its tables: each waits for the one before
    hipError_t e = hipStreamWaitEvent(st, B->done, 0);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "stream wait", e);
    // the tables staged by build_plan, once
    for (const BatchPlan::Upload &u : B->pending) {
        e = hipMemcpyAsync(u.d->ptr, static_cast<uint8_t *>(B->h_stage.ptr) + u.off, u.bytes,
                           hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "upload batch tables", e);
    }
    B->pending.clear();
    if (B->exec) {
        e = hipGraphLaunch(B->exec, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "batch graph launch", e);
    } else {
        int rc = B->record(st);
        if (rc != P2G_OK) return rc;
    }
    // error word, completion stamps and each group's role stamps to the host
    uint8_t *back = static_cast<uint8_t *>(B->h_back.ptr);
    e = hipMemcpyAsync(back, B->d_err.ptr, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(back + 8, B->d_done.ptr, (size_t)(B->njobs + 1) * 8, hipMemcpyDeviceToHost, st);
    for (size_t k = 0; e == hipSuccess && k < B->groups.size(); k++)
        e = hipMemcpyAsync(back + 8 + (size_t)(B->njobs + 1) * 8 + k * 8 * sizeof(unsigned long long),
                           static_cast<uint32_t *>(B->d_sync.ptr) + B->sync_off[k] + kPassTimes,
                           8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "copy batch results back", e);
    e = hipEventRecord(B->done, st);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "event record", e);
    B->launched = true;
    return P2G_OK;
}

int wait_plan(BatchPlan *B) {
    Ctx *c = B->c;
    hipSetDevice(c->device);
    if (!B->launched) return P2G_OK;
    hipError_t e = hipEventSynchronize(B->done);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "batch wait", e);
    B->launched = false;
    // every launch since the last wait ORs its errors into the sticky word;
    // the last launch copied it, its stamps and its role stamps to h_back
    const uint8_t *back = static_cast<const uint8_t *>(B->h_back.ptr);
    const unsigned long long *back_done = reinterpret_cast<const unsigned long long *>(back + 8);
    const unsigned long long *back_roles = back_done + B->njobs + 1;
    uint32_t bad = 0;
    memcpy(&bad, back, 4);
    if (bad) e = hipMemset(B->d_err.ptr, 0, 4);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "read batch flags", e);
    // the run's own role times refine the rates the next plans are sized by
    for (size_t k = 0; k < B->groups.size() && k < B->work.size(); k++) {
        const unsigned long long *t = back_roles + k * 8;
        auto span_us = [&](int role) {  // 100 MHz ticks
            const unsigned long long a = t[2 * role], b = t[2 * role + 1];
            return b > a && b != 0 ? (double)(b - a) / 100.0 : 0.0;
        };
        if (!bad) {
            const BatchGroupLaunch &g = B->groups[k];
            update_rates(c->rates, B->work[k], g.grid, g.h2_blocks, g.h3_blocks, span_us(kPassEncode),
                         span_us(kPassSlice));
        }
    }
    if (getenv("P2G_BATCH_VERBOSE")) {
        // per group: first start and last end of each role, ms from the batch start
        const unsigned long long t0[1] = {back_done[B->njobs]};
        for (size_t k = 0; k < B->groups.size(); k++) {
            const unsigned long long *t = back_roles + k * 8;
            const BatchGroupLaunch &g = B->groups[k];
            fprintf(stderr, "batch group %zu: pieces %u units %u tasks %u slices %u h3 %u h2 %u |", k,
                    g.n_pieces, g.n_units, g.n_tasks, g.slice_hi - g.slice_lo, g.h3_blocks, g.h2_blocks);
            static const char *names[4] = {"packet", "slice", "encode", "crc"};
            for (int r = 0; r < (int)kPassRoles; r++) {
                const unsigned long long a = t[2 * r], b = t[2 * r + 1];
                if (b == 0) continue;
                fprintf(stderr, " %s %.2f-%.2f", names[r], (double)(a - t0[0]) * 1e-5,
                        (double)(b - t0[0]) * 1e-5);
            }
            fprintf(stderr, "\n");
        }
    }
    B->h_done.assign(back_done, back_done + B->njobs + 1);
    if (bad & kPassErrWait)
        return c->fail(P2G_ERR_DEVICE,
                       "batched create: a packet hash stopped waiting for the encode (no progress for 2 s)");
    return P2G_OK;
}

int create_pass_batched(Ctx *c, const p2g_batch_job &job, hipStream_t st) {
    // keep at most two such plans (and their fragment buffers) alive for a
    // caller that queues many sets before synchronizing
    while (c->pending_batches.size() >= 2) {
        BatchPlan *old = c->pending_batches.front();
        c->pending_batches.erase(c->pending_batches.begin());
        const int rc = wait_plan(old);
        free_batch_plan(old);
        if (rc != P2G_OK) return rc;
    }
    BatchPlan *B = nullptr;
    int rc = make_plan(c, &job, 1, false, &B);
    if (rc != P2G_OK) return rc;
    rc = launch_plan(B, st);
    if (rc != P2G_OK) {
        delete B;
        return rc;
    }
    c->pending_batches.push_back(B);  // p2g_synchronize waits, checks and frees it
    return P2G_OK;
}

}  // namespace p2g

using namespace p2g;

extern "C" {

int p2g_batch_plan(p2g_ctx *ctx, const p2g_batch_job *jobs, uint32_t njobs, int use_graph,
                   p2g_batch **out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c || !out) return P2G_ERR_INVALID_ARGUMENT;
        if (const char *v = getenv("P2G_BATCH_GRAPH")) use_graph = atoi(v);
        BatchPlan *B = nullptr;
        int rc = make_plan(c, jobs, njobs, use_graph != 0, &B);
        *out = reinterpret_cast<p2g_batch *>(B);
        return rc;
    });
}

int p2g_batch_launch(p2g_batch *b, void *stream) {
    return guarded(b ? reinterpret_cast<BatchPlan *>(b)->c : nullptr, [&]() -> int {
        BatchPlan *B = reinterpret_cast<BatchPlan *>(b);
        if (!B) return P2G_ERR_INVALID_ARGUMENT;
        return launch_plan(B, stream ? static_cast<hipStream_t>(stream) : B->c->stream);
    });
}

int p2g_batch_wait(p2g_batch *b) {
    return guarded(b ? reinterpret_cast<BatchPlan *>(b)->c : nullptr, [&]() -> int {
        BatchPlan *B = reinterpret_cast<BatchPlan *>(b);
        if (!B) return P2G_ERR_INVALID_ARGUMENT;
        return wait_plan(B);
    });
}

int p2g_batch_job_ms(p2g_batch *b, double *ms_out) {
    return guarded(b ? reinterpret_cast<BatchPlan *>(b)->c : nullptr, [&]() -> int {
        BatchPlan *B = reinterpret_cast<BatchPlan *>(b);
        if (!B || !ms_out) return P2G_ERR_INVALID_ARGUMENT;
        if (B->h_done.size() != B->njobs + 1)
            return B->c->fail(P2G_ERR_INVALID_ARGUMENT, "no finished batch launch to report");
        const unsigned long long t0 = B->h_done[B->njobs];
        for (uint32_t j = 0; j < B->njobs; j++)
            ms_out[j] = B->h_done[j] >= t0 ? (double)(B->h_done[j] - t0) * 1e-5 : 0.0;
        return P2G_OK;
    });
}

void p2g_batch_free(p2g_batch *b) { delete reinterpret_cast<BatchPlan *>(b); }

int p2g_plan_roles(int kind, uint32_t num_cus, uint32_t max_resident, uint32_t slices,
                   uint32_t recovery_count, uint64_t slice_size, uint32_t *roles_out, uint32_t cap,
                   uint32_t *grid_out) {
    return guarded(nullptr, [&]() -> int {
        if (!grid_out || (cap && !roles_out) || slices == 0 || slices > kMaxInputs ||
            recovery_count == 0 || recovery_count > P2G_MAX_RECOVERY_BLOCKS || slice_size == 0 ||
            slice_size % 4 || slice_size > kMaxSliceBytes || num_cus == 0 || max_resident == 0)
            return P2G_ERR_INVALID_ARGUMENT;
        const uint32_t S = (uint32_t)slice_size, R = recovery_count, n = slices;
        uint32_t grid, h2, h3;
        if (kind == 0) {
            PassPlan p;
            if (!plan_create_pass(n, R, S, num_cus, max_resident, 0, &p)) return P2G_ERR_INVALID_ARGUMENT;
            grid = p.grid;
            h2 = p.h2_blocks;
            h3 = p.h3_blocks;
        } else if (kind == 1) {
            // a one-job plan with one piece, as build_plan sizes it
            BatchPiece P{};
            P.n = n;
            P.nq = mfma_steps(n);
            P.R = R;
            P.S = S;
            P.row_blocks = (R + 15) / 16;
            P.col_groups = P.cg_total = (S + 511) / 512;
            BatchGroupWork w;
            add_slice_work(w, n, S);
            add_piece_work(w, P);
            BatchGroupLaunch g{};
            const CostRates seed;
            plan_batch_group(num_cus, max_resident, (R + 63) / 64, n, P.row_blocks * P.col_groups, w,
                             seed, &g);
            grid = g.grid;
            h2 = g.h2_blocks;
            h3 = g.h3_blocks;
        } else {
            return P2G_ERR_INVALID_ARGUMENT;
        }
        *grid_out = grid;
        for (uint32_t b = 0; b < grid && b < cap; b++) roles_out[b] = pass_role(b, grid, h2, h3).role;
        return P2G_OK;
    });
}

int p2g_dev_create_batch(p2g_ctx *ctx, const p2g_batch_job *jobs, uint32_t njobs, void *stream) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        BatchPlan *B = nullptr;
        int rc = make_plan(c, jobs, njobs, false, &B);
        if (rc != P2G_OK) return rc;
        rc = launch_plan(B, stream ? static_cast<hipStream_t>(stream) : c->stream);
        if (rc != P2G_OK) {
            delete B;
            return rc;
        }
        c->pending_batches.push_back(B);  // p2g_synchronize waits, checks and frees it
        return P2G_OK;
    });
}

}  // extern "C"
