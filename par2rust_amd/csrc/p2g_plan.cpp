// p2g_plan.cpp -- host-side planning of the persistent create launches:
// the role split of the fused pass (plan_create_pass) and of a batched
// launch group (plan_batch_group).  Host only: no device code and no HIP
// calls, so it also builds with a plain C++ compiler (the sanitizer harness
// does).
#include <algorithm>
#include <cmath>
#include <vector>

#include "p2g_internal.hpp"

namespace p2g {

bool plan_create_pass(uint32_t n, uint32_t R, uint32_t S, uint32_t num_cus, uint32_t max_resident,
                      uint32_t band_cols, PassPlan *p) {
    if (n == 0 || R == 0 || 4 * mfma_steps(n) > (uint32_t)kMfMaxInputs) return false;
    // one workgroup per CU, and never more than the device holds at once
    // (the roles wait on each other across workgroups)
    p->grid = std::min(num_cus, max_resident);
    // one 64-lane wave per MD5 workgroup; lanes loop when the caps bind
    p->h3_blocks = std::min<uint32_t>((R + 63) / 64, 8);
    p->h2_blocks = std::min<uint32_t>((n + 63) / 64, 32);
    if (4 * (p->h2_blocks + p->h3_blocks) >= p->grid) return false;
    p->row_blocks = mfma_row_tiles(R) / kMfTiles;
    p->col_groups = (S + kMfThreads * 2 - 1) / (kMfThreads * 2);
    p->band_cols = std::max<uint32_t>(1, std::min(band_cols ? band_cols : kPassDefaultBand, p->col_groups));
    p->n_bands = 0;
    for (uint32_t c = 0; c < p->col_groups; p->n_bands++)
        c += next_band_width(c, p->col_groups - c, p->band_cols);
    p->n_units = p->row_blocks * p->col_groups;
    return true;
}

// A measurement is taken when the role ran long enough to be timed well,
// kept within [1/2, 2] x the seed value (a run sharing the GPU with other
// work must not swing the sizing), and averaged in with weight 1/2.
static void fold(double &rate, double seed, double meas) {
    if (!(meas > 0)) return;
    meas = std::min(std::max(meas, 0.5 * seed), 2.0 * seed);
    rate = 0.5 * rate + 0.5 * meas;
}

void update_rates(CostRates &r, const BatchGroupWork &w, uint32_t grid, uint32_t h2, uint32_t h3,
                  double enc_us, double slice_us) {
    const CostRates seed;
    bool used = false;
    const uint32_t e = grid > h2 + h3 ? grid - h2 - h3 : 1;
    if (enc_us > 500 && w.enc_steps > 0) {  // the encode: e workgroups over its unit-steps
        fold(r.step_us, seed.step_us, enc_us * e / w.enc_steps);
        used = true;
    }
    if (slice_us > 500 && h2 && w.slice_blocks > 0) {  // the slice lanes, net of the per-slice cost
        const double lanes = (double)kBatchSliceLanes * h2;
        fold(r.slice_block_us, seed.slice_block_us,
             (slice_us * lanes - w.slices * r.slice_us) / w.slice_blocks);
        used = true;
    }
    if (used) r.updates++;
}


// Estimated end (µs) of the packet role on 4 h3 waves when the encode takes
// t_enc: pieces are encoded in order, each over its share of t_enc by its
// unit-steps; task t runs on wave t mod 4 h3 after the wave's earlier tasks,
// starts no earlier than its piece's encode and, following its bands, ends
// no earlier than the piece's encode ends.
static double packet_role_end(const BatchGroupWork &w, uint32_t h3, double t_enc,
                              const CostRates &rt) {
    if (h3 == 0 || w.enc_steps <= 0) return 0;
    std::vector<double> free_at(4 * h3, 0.0);
    double done = 0, end = 0;
    size_t t = 0;
    for (size_t p = 0; p < w.piece_steps.size(); p++) {
        const double start = t_enc * done / w.enc_steps;
        done += w.piece_steps[p];
        const double finish = t_enc * done / w.enc_steps;
        const double chain = rt.task_us + w.piece_task_blocks[p] * rt.packet_block_us;
        for (double k = 0; k < w.piece_tasks[p]; k++, t++) {
            double &f = free_at[t % free_at.size()];
            f = std::max(std::max(f, start) + chain, finish);
            end = std::max(end, f);
        }
    }
    return end;
}

void plan_batch_group(uint32_t num_cus, uint32_t max_resident, uint32_t n_tasks, uint32_t n_slices,
                      uint32_t n_units, const BatchGroupWork &w, const CostRates &rt,
                      BatchGroupLaunch *g) {
    // one workgroup per CU, never more than the device holds at once
    num_cus = std::max<uint32_t>(1, std::min(num_cus, max_resident));
    g->grid = num_cus;
    // four packet waves per packet workgroup, kBatchSliceLanes MD5 + CRC-32
    // lanes per slice workgroup; the rest encode
    const uint32_t h3_cap = std::min<uint32_t>((n_tasks + 3) / 4, 8);
    const uint32_t h2_cap = std::min<uint32_t>((n_slices + kBatchSliceLanes - 1) / kBatchSliceLanes, 64);
    if (n_units == 0) {
        g->h3_blocks = h3_cap;
        g->h2_blocks = h2_cap;
        g->grid = std::max<uint32_t>(1, g->h2_blocks + g->h3_blocks);
        return;
    }
    // A hash workgroup holds its CU until its role is done, so each role
    // gets the fewest workgroups that keep it within the encode on the
    // remaining CUs.  The slice waves take 64-slice tasks from a queue that
    // runs from the largest slices to the smallest, so the slice lanes'
    // estimate is their average work plus one slice of the smallest size,
    // kept within 3/4 of the encode.  The packet tasks follow
    // the encode band by band: workgroups are added while the simulated end
    // of their role lies more than 2% past the encode and more workgroups
    // would bring it closer.
    const double slice_us = w.slices * rt.slice_us + w.slice_blocks * rt.slice_block_us;
    uint32_t h3 = std::min<uint32_t>(h3_cap, 1), h2 = std::min<uint32_t>(h2_cap, 1);
    for (;;) {
        const uint32_t e = num_cus > h2 + h3 + 1 ? num_cus - h2 - h3 : 1;
        const double t_enc = w.enc_steps * rt.step_us / e;
        bool grew = false;
        const double tail_us = rt.slice_us + w.small_blocks * rt.slice_block_us;
        if (h2 < h2_cap && slice_us / ((double)kBatchSliceLanes * h2) + tail_us > 0.75 * t_enc) {
            h2++;
            grew = true;
        }
        const double end = packet_role_end(w, h3, t_enc, rt);
        if (h3 < h3_cap && end > 1.02 * t_enc) {
            // the fewest more workgroups that bring the end closer (waves
            // that still hold two tasks of one piece gain nothing)
            for (uint32_t h = h3 + 1; h <= h3_cap; h++) {
                if (packet_role_end(w, h, t_enc, rt) < end - 0.01 * t_enc) {
                    h3 = h;
                    grew = true;
                    break;
                }
            }
        }
        if (!grew) break;
    }
    g->h3_blocks = h3;
    g->h2_blocks = h2;
    if (g->h2_blocks + g->h3_blocks + 1 > num_cus) {
        g->h3_blocks = std::min<uint32_t>(g->h3_blocks, 1);
        g->h2_blocks = std::min<uint32_t>(g->h2_blocks, 1);
    }
    // at least one encode workgroup below the hash roles
    g->grid = std::max(g->grid, g->h2_blocks + g->h3_blocks + 1);
}

}  // namespace p2g
