// Host-only harness for the planners of the persistent create launches
// (p2g_plan.cpp) and the volume split (par2_scheme.cpp), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitizers.py.  Sweeps shapes and schemes, checks each
// result's invariants, and exits non-zero on the first violation.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "p2g_internal.hpp"
#include "par2_create.hpp"

using namespace p2g;

static int fails = 0;
#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c);      \
            if (++fails > 20) exit(1);                                           \
        }                                                                        \
    } while (0)

static void check_roles(uint32_t grid, uint32_t h2, uint32_t h3) {
    CHECK(h2 + h3 < grid);
    uint32_t last = kPassEncode, seen[3] = {0, 0, 0};
    for (uint32_t b = 0; b < grid; b++) {
        const PassRole r = pass_role(b, grid, h2, h3);
        CHECK(r.role <= 2);
        seen[r.role]++;
        // encode, then slice, then packet: the role id only decreases
        CHECK(r.role <= last);
        last = r.role;
    }
    CHECK(seen[kPassEncode] == grid - h2 - h3 && seen[kPassSlice] == h2 && seen[kPassPacket] == h3);
}

static uint32_t lcg(uint64_t &s) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(s >> 33);
}

int main() {
    uint64_t rng = 12345;
    const uint32_t cus_list[] = {1, 2, 3, 8, 16, 64, 100, 256, 304};
    // fused pass
    uint32_t planned = 0;
    for (int t = 0; t < 20000; t++) {
        const uint32_t n = 1 + lcg(rng) % (t % 3 ? 2100 : 40000);
        const uint32_t R = 1 + lcg(rng) % (t % 5 ? 300 : 65535);
        const uint32_t S = 4 * (1 + lcg(rng) % (t % 7 ? 4096 : (1u << 29)));
        const uint32_t cus = cus_list[lcg(rng) % 9];
        const uint32_t resident = t % 4 ? cus : 1 + lcg(rng) % (2 * cus);
        const uint32_t band = t % 2 ? 0 : lcg(rng) % 300;
        PassPlan p;
        if (!plan_create_pass(n, R, S, cus, resident, band, &p)) continue;
        planned++;
        CHECK(p.grid <= cus && p.grid <= resident);
        CHECK(4 * mfma_steps(n) <= (uint32_t)kMfMaxInputs);
        check_roles(p.grid, p.h2_blocks, p.h3_blocks);
        CHECK(p.row_blocks * 16 >= R && (p.row_blocks - 1) * 16 < R);
        CHECK(p.col_groups * 512ull >= S && (p.col_groups - 1) * 512ull < S);
        uint32_t c = 0, bands = 0;
        while (c < p.col_groups) {
            const uint32_t w = next_band_width(c, p.col_groups - c, p.band_cols);
            CHECK(w >= 1 && w <= p.band_cols);
            c += w;
            bands++;
            if (bands > p.col_groups) break;
        }
        CHECK(c == p.col_groups && bands == p.n_bands);
        CHECK(p.n_units == p.row_blocks * p.col_groups);
    }
    CHECK(planned > 1000);
    // the fused pass's unit order inside a band: every (row block, column
    // group) of the band exactly once, for any tile width
    for (uint32_t rows = 1; rows <= 40; rows += 3)
        for (uint32_t w = 1; w <= 70; w += 7)
            for (uint32_t tc : {1u, 2u, 8u, 9u, 33u}) {
                std::vector<uint8_t> seen(rows * w, 0);
                for (uint32_t v = 0; v < rows * w; v++) {
                    uint32_t y = 0, x = 0;
                    band_tile_unit(v, w, rows, tc, &y, &x);
                    CHECK(y < rows && x < w);
                    if (y < rows && x < w) seen[y * w + x]++;
                }
                for (uint8_t k : seen) CHECK(k == 1);
            }
    // batched groups
    for (int t = 0; t < 5000; t++) {
        BatchGroupWork w;
        const uint32_t pieces = 1 + lcg(rng) % 20;
        uint32_t tasks = 0, units = 0, slices = 0;
        for (uint32_t k = 0; k < pieces; k++) {
            const uint32_t n = 1 + lcg(rng) % 32768, R = 1 + lcg(rng) % 3000;
            const uint32_t S = 4 * (1 + lcg(rng) % (1u << 18));
            const uint32_t rb = (R + 15) / 16, cg = (S + 511) / 512;
            const double steps = (double)rb * cg * mfma_steps(n);
            w.enc_steps += steps;
            w.piece_steps.push_back(steps);
            w.piece_tasks.push_back((R + 63) / 64);
            w.piece_task_blocks.push_back((double)((36ull + S + 8) / 64 + 1));
            tasks += (R + 63) / 64;
            units += t % 11 ? rb * cg : 0;
            slices += n;
            const double blocks = (double)((S + 8) / 64 + 1);
            w.slices += n;
            w.slice_blocks += n * blocks;
            if (blocks > w.big_blocks) {
                w.big_blocks = blocks;
                w.big_slices = 0;
            }
            if (blocks == w.big_blocks) w.big_slices += n;
            if (w.small_blocks == 0 || blocks < w.small_blocks) w.small_blocks = blocks;
        }
        const uint32_t cus = cus_list[lcg(rng) % 9];
        const uint32_t resident = t % 3 ? cus : 1 + lcg(rng) % (2 * cus);
        BatchGroupLaunch g{};
        CostRates rates;
        if (t % 2) {  // rates a run measured, within the range update_rates keeps
            update_rates(rates, w, cus, 2, 1, 1000.0 + lcg(rng) % 100000, 1000.0 + lcg(rng) % 50000);
            CHECK(rates.step_us >= 0.19 && rates.step_us <= 0.76);
            CHECK(rates.slice_block_us >= 0.65 && rates.slice_block_us <= 2.62);
        }
        plan_batch_group(cus, resident, tasks, slices, units, w, rates, &g);
        CHECK(g.h3_blocks <= 8 && g.h2_blocks <= 64);
        if (units) {
            check_roles(g.grid, g.h2_blocks, g.h3_blocks);
            CHECK(g.grid <= std::max<uint32_t>(std::min(cus, resident), g.h2_blocks + g.h3_blocks + 1));
            CHECK(g.h3_blocks >= 1 && g.h2_blocks >= 1);
        }
    }
    // volume schemes
    for (int t = 0; t < 20000; t++) {
        par2::VolumeScheme s;
        s.kind = (int)(lcg(rng) % 5);  // 4 is not a scheme
        const uint32_t total = lcg(rng) % (t % 10 ? 3000 : 65536);
        s.uniform_count = lcg(rng) % 40;
        const uint32_t ncaps = lcg(rng) % 4;
        for (uint32_t k = 0; k < ncaps; k++) s.caps.push_back(lcg(rng) % 50);
        if (s.kind == P2G_SCHEME_EXPLICIT) {
            uint32_t left = total;
            const uint32_t parts = lcg(rng) % 6;
            for (uint32_t k = 0; k < parts; k++) {
                const uint32_t v = k + 1 == parts ? left : (left ? lcg(rng) % (left + 1) : 0);
                s.explicit_sizes.push_back(v);
                left -= v;
            }
        }
        std::vector<uint32_t> out;
        par2::Error err;
        const int rc = par2::resolve_volume_sizes(s, total, out, err);
        if (rc != P2G_OK) {
            CHECK(rc == P2G_ERR_INVALID_VOLUME_SCHEME && !err.message.empty());
            continue;
        }
        uint64_t sum = 0;
        for (uint32_t v : out) {
            CHECK(v > 0 || total == 0);
            for (uint32_t cap : s.caps) CHECK(v <= cap);  // each cap splits, none merges
            sum += v;
        }
        CHECK(sum == total);
    }
    // volume names
    CHECK(par2::volume_name("a/b/recovery.par2", 0, 4) == "a/b/recovery.vol0+4.par2");
    CHECK(par2::volume_name("noext", 3, 1) == "noext.vol3+1.par2");
    CHECK(par2::volume_name(".hidden", 1, 2) == ".hidden.vol1+2.par2");
    CHECK(par2::volume_name("x.tar.gz", 65534, 1) == "x.tar.vol65534+1.gz");
    if (fails) return 1;
    printf("ok %u fused plans\n", planned);
    return 0;
}
