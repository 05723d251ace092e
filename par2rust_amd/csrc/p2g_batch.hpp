// p2g_batch.hpp -- the batched create plan (p2g_batch.cpp), shared with the
// serving loop (p2g_serve.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "p2g_context.hpp"

namespace p2g {

// Tables and launch sequence of one batched run: descriptors of every
// item's slices (partial slices redirected to zero-padded staging copies),
// pieces, slice jobs, and per launch group its sync words; the A fragments
// live in `afrag`, the plan's own buffer or one shared by plans whose runs
// are ordered on one stream.
struct BatchPlan {
    Ctx *c = nullptr;
    uint32_t njobs = 0;  // items: completion stamps are per item
    uint32_t n_slices = 0, n_parts = 0, max_part_S = 0;
    DevBuf d_descs, d_tb, d_staging, d_parts, d_pieces, d_sjobs, d_sync, d_done;
    DevBuf d_err;  // sticky error word of every launch of the plan, cleared by wait_plan
    DevBuf own_afrag;
    DevBuf *afrag = &own_afrag;  // or a buffer shared by plans run in one stream order
    std::vector<BatchGroupLaunch> groups;
    std::vector<BatchGroupWork> work;  // each group's work, for the rates its run measures
    std::vector<uint32_t> sync_off;  // word offset of each group's sync words
    // Host side of the plan's transfers, all asynchronous on the launch's
    // stream (pinned memory): the tables are staged here by build_plan and
    // uploaded by the next launch; each launch copies its error word,
    // completion stamps and role stamps back before its `done` event.  No
    // host call then waits on the device while another plan's persistent
    // launch, which holds every CU, is running.
    struct Upload {
        DevBuf *d;
        size_t off, bytes;
    };
    PinBuf h_stage, h_back;
    std::vector<Upload> pending;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
    std::vector<unsigned long long> h_done;  // completion stamps of the last launch, then start
    // A fragments of one group, in bytes; 0: a quarter of the free device
    // memory within [16, 64] GiB (P2G_BATCH_AFRAG_BYTES overrides either).
    uint64_t afrag_budget = 0;
    // Freeing device or pinned memory waits for every kernel on the device.
    // A plan rebuilt while another plan's launch runs (the serving steps)
    // sets keep_retired: a table that outgrows its buffer then moves to a new
    // one and the old one waits here, freed by free_retired (when the caller
    // knows the device idle) or with the plan.
    bool keep_retired = false;
    std::vector<void *> retired_dev, retired_host;
    uint64_t retired_bytes = 0;

    int record(hipStream_t st);
    void drop_graph();
    void free_retired();
    ~BatchPlan();
};

}  // namespace p2g
