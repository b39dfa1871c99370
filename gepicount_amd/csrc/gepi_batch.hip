/*
 * gepi_batch.hip -- the batched scan engine (Engine) and its C ABI
 * (gepi_engine_*).  Kernels are in gepi_batch.h; the per-scan band logic is
 * the one of run_scan (gepi_scan.hip) applied to every active scan at once.
 * No counterpart in the reference, which runs one scan per process
 * (src/gEpiCount.cpp:84-87).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/gepicount.h"
#include "gepi_batch.h"
#include "gepi_engine.h"
#include "gepi_host.h"

using namespace gepi;

namespace gepi_host {

namespace {

constexpr int64_t F4_TILE_EVALS = (int64_t)TM * TB4 * 20;
constexpr int EXACT_BLOCKS_X = 64; /* blocks per scan of the batched exact pass */

/* Job shape of an entropy scan with ks K steps under the default schedule:
 * wave tiles from 48 K steps, whole tiles below.  Measured on the MI355X
 * with 4 scans of 12,000 variants per batch against whole tiles in the
 * persistent launch, count launches: 1.27-1.32x the time at 1,536-2,048
 * samples, 1.10-1.14x at 3,072-4,096, 0.97x at 6,144, 0.85x at 8,192.     */
constexpr int32_t BATCH_WAVE_MIN_KSTEPS = 48;
int batch_shape_default(int32_t ks) { return ks >= BATCH_WAVE_MIN_KSTEPS ? 2 : 0; }

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

size_t round_block(size_t bytes) {
    /* 64 KiB granules below 64 MiB, 1 MiB above: cached blocks are reused for
     * any request between half and all of their size                       */
    const size_t g = bytes < ((size_t)64 << 20) ? ((size_t)64 << 10) : ((size_t)1 << 20);
    return std::max<size_t>(g, (bytes + g - 1) / g * g);
}

unsigned long long env_ull(const char *name, unsigned long long dflt) {
    const char *v = getenv(name);
    if (!v) return dflt;
    const long long x = atoll(v);
    return x > 0 ? (unsigned long long)x : dflt;
}

} // namespace

int Engine::init(int dev) {
    device = dev;
    int rc;
    if ((rc = set_device(dev)) != GEPI_OK) return rc;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&ev0_));
    HIP_TRY(hipEventCreate(&ev1_));
    /* fp4 work counters: per objective launch, 2 sets of 8 */
    HIP_TRY(hipMalloc(&d_wq_, sizeof(uint32_t) * 16 * NGROUP));
    HIP_TRY(hipMemsetAsync(d_wq_, 0, sizeof(uint32_t) * 16 * NGROUP, stream));
    /* rounds as captured graphs only on request: on the MI355X a graph
     * launch costs about 0.16 ms per round more than the direct launches */
    const char *g = getenv("GEPI_ENGINE_GRAPH");
    use_graph_ = g && atoi(g) != 0;
    return GEPI_OK;
}

void Engine::drop_graphs() {
    for (auto &g : gx_count_) {
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
    if (gx_finish_) hipGraphExecDestroy(gx_finish_);
    gx_finish_ = nullptr;
    gx_cap_ = 0;
    gx_per_cu_ = 0;
}

void Engine::launch_count(int g, bool persist, unsigned grid, const BatchMap &m) {
    const BatchDesc *d = d_bdesc_[g];
    if (g >= 4) {
        /* wave tiles: one unit per workgroup, no persistent form */
        if (g == 4)
            k_count_f4w_batch<GEPI_FFUNC_ENTROPY><<<grid, F4W_THREADS, 0, stream>>>(d, m);
        else
            k_count_f4w_batch<GEPI_FFUNC_FISHERPV><<<grid, F4W_THREADS, 0, stream>>>(d, m);
        return;
    }
    switch (g | (persist ? 4 : 0)) {
    case 0: k_count_f4_batch<GEPI_FFUNC_ENTROPY, false, false><<<grid, F4_THREADS, 0, stream>>>(d, m); break;
    case 1: k_count_f4_batch<GEPI_FFUNC_FISHERPV, false, false><<<grid, F4_THREADS, 0, stream>>>(d, m); break;
    case 2: k_count_f4_batch<GEPI_FFUNC_ENTROPY, false, true><<<grid, F4_THREADS, 0, stream>>>(d, m); break;
    case 3: k_count_f4_batch<GEPI_FFUNC_FISHERPV, false, true><<<grid, F4_THREADS, 0, stream>>>(d, m); break;
    case 4: k_count_f4_batch<GEPI_FFUNC_ENTROPY, true, false><<<grid, F4_THREADS, 0, stream>>>(d, m); break;
    case 5: k_count_f4_batch<GEPI_FFUNC_FISHERPV, true, false><<<grid, F4_THREADS, 0, stream>>>(d, m); break;
    case 6: k_count_f4_batch<GEPI_FFUNC_ENTROPY, true, true><<<grid, F4_THREADS, 0, stream>>>(d, m); break;
    default: k_count_f4_batch<GEPI_FFUNC_FISHERPV, true, true><<<grid, F4_THREADS, 0, stream>>>(d, m); break;
    }
}

int Engine::capture_graph(int part, int per_cu) {
    const size_t cap = desc_cap_;
    const unsigned grid =
        (unsigned)std::max<int64_t>(8, (per_cu * (int64_t)device_cus(device)) / 8 * 8);
    /* nothing may return between the begin and the end of the capture */
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    if (part < NGRAPH_COUNT) {
        k_fill_u32<<<1, 64, 0, stream>>>(d_wq_, 16 * NGROUP, 0u);
        for (int g = 0; g < 4; g++) {
            if (!(part & (1 << g))) continue;
            /* the maps are laid out for cap scans; the slots past a round's
             * scans have empty ranges                                       */
            BatchMap m;
            m.xprefix = d_map_[g];
            m.xstart = d_map_[g] + 8 * (cap + 1);
            m.n = (int32_t)cap;
            m.wq = d_wq_ + 16 * g;
            m.wq_parity = 0;
            launch_count(g, true, grid, m);
        }
    } else {
        k_exact_batch<<<dim3(EXACT_BLOCKS_X, (unsigned)cap), 256, 0, stream>>>(d_edesc_);
        k_topk_wg<<<(unsigned)cap, TOPK_THREADS, 0, stream>>>(d_tdesc_);
    }
    const hipError_t le = hipGetLastError();
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(stream, &graph);
    hipError_t ie = hipSuccess;
    if (le == hipSuccess && ce == hipSuccess)
        ie = hipGraphInstantiate(part < NGRAPH_COUNT ? &gx_count_[part] : &gx_finish_, graph, nullptr,
                                 nullptr, 0);
    if (graph) hipGraphDestroy(graph);
    if (le != hipSuccess || ce != hipSuccess || ie != hipSuccess) {
        const hipError_t e = le != hipSuccess ? le : ce != hipSuccess ? ce : ie;
        drop_graphs();
        return gepi_fail(GEPI_ERR_DEVICE, "round graph capture: %s", hipGetErrorString(e));
    }
    return GEPI_OK;
}

void Engine::release() {
    /* idempotent: every handle is cleared after it is released */
    if (stream) {
        hipSetDevice(device);
        hipStreamSynchronize(stream);
    }
    drop_graphs();
    for (ScanItem *it : active_) drop(it, GEPI_ERR_INTERNAL);
    active_.clear();
    for (auto &b : free_) hipFree(b.second);
    free_.clear();
    for (auto &b : live_) hipFree(b.first);
    live_.clear();
    in_use_ = 0;
    for (int i = 0; i < NGROUP; i++) {
        hipFree(d_bdesc_[i]);
        hipFree(d_map_[i]);
        d_bdesc_[i] = nullptr;
        d_map_[i] = nullptr;
    }
    hipFree(d_edesc_);
    hipFree(d_tdesc_);
    hipFree(d_rep_);
    hipFree(d_wq_);
    d_wq_ = nullptr;
    hipFree(d_park_);
    d_park_ = nullptr;
    park_slots_ = 0;
    d_edesc_ = nullptr;
    d_tdesc_ = nullptr;
    d_rep_ = nullptr;
    desc_cap_ = 0;
    hipFree(tabs_.lhi);
    hipFree(tabs_.llo);
    hipFree(tabs_.lfact);
    tabs_ = DeviceTables();
    if (ev0_) hipEventDestroy(ev0_);
    if (ev1_) hipEventDestroy(ev1_);
    ev0_ = ev1_ = nullptr;
    if (stream) hipStreamDestroy(stream);
    stream = nullptr;
}

void *Engine::dalloc(size_t bytes) {
    const size_t sz = round_block(bytes);
    void *p = nullptr;
    auto it = free_.lower_bound(sz);
    if (it != free_.end() && it->first <= 2 * sz) {
        p = it->second;
        live_[p] = it->first;
        in_use_ += it->first;
        free_.erase(it);
        return p;
    }
    if (hipMalloc(&p, sz) != hipSuccess) {
        /* return the cache to the device and try once more */
        for (auto &b : free_) hipFree(b.second);
        free_.clear();
        if (hipMalloc(&p, sz) != hipSuccess) return nullptr;
    }
    live_[p] = sz;
    in_use_ += sz;
    return p;
}

void Engine::dfree(void *p) {
    if (!p) return;
    auto it = live_.find(p);
    if (it == live_.end()) return;
    in_use_ -= it->second;
    free_.insert({it->second, p});
    live_.erase(it);
}

const gepi_tune_class *Engine::tune_of(int32_t n_k4) const {
    if (tuning.empty()) return nullptr;
    for (const gepi_tune_class &c : tuning)
        if (n_k4 <= c.max_ksteps) return &c;
    return &tuning.back();
}

int Engine::time_batch(std::vector<ScanItem> &items, double *ms) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    for (ScanItem &it : items) add(&it);
    std::vector<ScanItem *> fin;
    while (!idle())
        if ((rc = round(fin)) != GEPI_OK) {
            /* the items are the caller's: none may stay active */
            for (ScanItem *it : active_) drop(it, rc);
            active_.clear();
            return rc;
        }
    *ms = ms_since(t0);
    for (const ScanItem &it : items)
        if (it.status != GEPI_OK) return it.status;
    return GEPI_OK;
}

namespace {

/* Synthetic SNP-major .bed bytes: per variant an allele frequency in
 * [0.05, 0.5], genotypes by Hardy-Weinberg, 1% missing calls (xorshift64*). */
std::vector<uint8_t> synth_bed(int64_t n_var, int32_t n, uint64_t seed) {
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&s]() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 2685821657736338717ull;
    };
    auto unif = [&next]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); };
    const int64_t bpv = (n + 3) / 4;
    std::vector<uint8_t> bed((size_t)(n_var * bpv), 0);
    for (int64_t v = 0; v < n_var; v++) {
        const double f = 0.05 + 0.45 * unif();
        const double p0 = (1 - f) * (1 - f), p1 = p0 + 2 * f * (1 - f);
        for (int32_t i = 0; i < n; i++) {
            const double u = unif();
            /* .bed codes: 0 homozygous A1, 2 heterozygous, 3 homozygous A2, 1 missing */
            const uint8_t code = unif() < 0.01 ? 1 : u < p0 ? 0 : u < p1 ? 2 : 3;
            bed[(size_t)(v * bpv + i / 4)] |= (uint8_t)(code << (2 * (i % 4)));
        }
    }
    return bed;
}

} // namespace

int Engine::autotune(int32_t repeats) {
    struct Trial {
        int32_t max_k4;
        int64_t n_var;
        int32_t n;
    };
    /* trial batches large enough that a schedule's per-job costs, not the
     * tail of a short launch, decide (tens of ms each) */
    const Trial trials[4] = {{2, 40000, 256}, {8, 28000, 1024}, {32, 18000, 4096}, {INT32_MAX, 12000, 8192}};
    const int per_cus[4] = {0, 1, 2, 3};
    const double bands[3] = {5e8, 2e9, 8e9};
    /* job shapes: tiles, half tiles, wave tiles (not in captured rounds) */
    const int n_tiles = use_graph_ ? 2 : 3;
    const int tiles[3] = {0, 1, 2};
    /* the default schedule's shape for a class's trial scans (about half
     * cases); captured rounds run whole tiles in place of wave tiles     */
    auto dflt_tile = [&](int c) {
        return use_graph_ ? 0 : batch_shape_default((trials[c].n / 2 + 127) / 128 * 2);
    };
    repeats = std::max(1, std::min(repeats, 16));
    std::vector<gepi_tune_class> table;
    int rc = GEPI_OK;
    tuning.clear();
    /* every class's panel and 4 phenotypes stay until the end: the table is
     * checked on all 16 scans together */
    std::vector<gepi_panel *> panels(4, nullptr);
    std::vector<std::vector<uint64_t>> masks;
    std::vector<ScanItem> all;
    auto scans_of = [&](int c) {
        std::vector<ScanItem> items(4);
        for (int j = 0; j < 4; j++) {
            items[(size_t)j].panel = panels[(size_t)c];
            items[(size_t)j].mask_ctrl = masks[(size_t)(8 * c + 2 * j)].data();
            items[(size_t)j].mask_case = masks[(size_t)(8 * c + 2 * j + 1)].data();
            items[(size_t)j].topmost = 100;
            items[(size_t)j].ffun = GEPI_FFUNC_ENTROPY;
        }
        return items;
    };
    for (int c = 0; c < 4 && rc == GEPI_OK; c++) {
        const Trial &t = trials[c];
        const std::vector<uint8_t> bed = synth_bed(t.n_var, t.n, 1000 + c);
        if ((rc = gepi_panel_create_bed(bed.data(), t.n_var, t.n, device, &panels[(size_t)c])) !=
            GEPI_OK)
            break;
        /* 4 phenotypes of about half cases each */
        const int32_t w64 = (t.n + 63) / 64;
        uint64_t s = 77 + (uint64_t)c;
        for (int j = 0; j < 4; j++) {
            std::vector<uint64_t> m0((size_t)w64, 0), m1((size_t)w64, 0);
            for (int32_t i = 0; i < t.n; i++) {
                s = s * 6364136223846793005ull + 1442695040888963407ull;
                ((s >> 40) & 1 ? m1 : m0)[(size_t)(i / 64)] |= 1ull << (i % 64);
            }
            masks.push_back(std::move(m0));
            masks.push_back(std::move(m1));
        }
        std::vector<ScanItem> items = scans_of(c);
        const int dtile = dflt_tile(c);
        gepi_tune_class best{t.max_k4, 2, dtile, 0, 2e9, 0.0, 0.0};
        double ms_best = INFINITY, ms_dflt = INFINITY, ms;
        /* warm-up (allocations, tables, graphs) under the default, then
         * once per job shape (the first launches of its kernels) */
        for (int ti = 0; ti < n_tiles; ti++) {
            tuning.assign(1, gepi_tune_class{INT32_MAX, 2, tiles[ti], 0, 2e9, 0.0, 0.0});
            if (rc == GEPI_OK) rc = time_batch(items, &ms);
        }
        for (int r = 0; r < repeats && rc == GEPI_OK; r++)
            for (int ti = 0; ti < n_tiles; ti++)
                for (int pc : per_cus)
                    for (double b : bands) {
                        const int tl = tiles[ti];
                        /* wave tiles have one launch form (a unit per
                         * workgroup); their class keeps the default 2 for
                         * the rounds it leads                                */
                        if (tl == 2 && pc != 2) continue;
                        if (rc != GEPI_OK) break;
                        tuning.assign(1, gepi_tune_class{INT32_MAX, pc, tl, 0, b, 0.0, 0.0});
                        if ((rc = time_batch(items, &ms)) != GEPI_OK) break;
                        if (tl == dtile && pc == 2 && b == 2e9) ms_dflt = std::min(ms_dflt, ms);
                        /* the best of each schedule over the repeats */
                        if (ms < ms_best) {
                            ms_best = ms;
                            best.wg_per_cu = pc;
                            best.tile = tl;
                            best.band_evals = b;
                        }
                    }
        if (rc != GEPI_OK) break;
        if (ms_dflt <= 1.03 * ms_best) {
            /* within the trial's noise: keep the default */
            best.wg_per_cu = 2;
            best.tile = dtile;
            best.band_evals = 2e9;
            ms_best = ms_dflt;
        }
        best.ms_best = ms_best;
        best.ms_default = ms_dflt;
        table.push_back(best);
    }
    /* The classes interact in a mixed batch (a round's launch form follows
     * its largest class; each job shape present is a launch of its own), so
     * the table is kept only if it also beats the default table by 1% on
     * all 16 scans together; otherwise every class gets the default.      */
    if (rc == GEPI_OK) {
        for (int c = 0; c < 4; c++) {
            std::vector<ScanItem> items = scans_of(c);
            all.insert(all.end(), items.begin(), items.end());
        }
        bool differs = false;
        for (size_t c = 0; c < table.size(); c++)
            differs |= table[c].wg_per_cu != 2 || table[c].band_evals != 2e9 ||
                       table[c].tile != dflt_tile((int)c);
        double ms_t = INFINITY, ms_d = INFINITY, ms;
        for (int r = 0; r < repeats && rc == GEPI_OK && differs; r++) {
            tuning = table;
            if ((rc = time_batch(all, &ms)) != GEPI_OK) break;
            ms_t = std::min(ms_t, ms);
            tuning.clear();
            if ((rc = time_batch(all, &ms)) != GEPI_OK) break;
            ms_d = std::min(ms_d, ms);
        }
        if (rc == GEPI_OK && differs && !(ms_t < 0.99 * ms_d))
            for (size_t c = 0; c < table.size(); c++) {
                table[c].wg_per_cu = 2;
                table[c].tile = dflt_tile((int)c);
                table[c].band_evals = 2e9;
                table[c].ms_best = table[c].ms_default;
            }
        mixed_ms_tuned_ = ms_t;
        mixed_ms_default_ = ms_d;
    }
    for (gepi_panel *p : panels)
        if (p) gepi_panel_destroy(p);
    tuning = rc == GEPI_OK ? table : std::vector<gepi_tune_class>();
    return rc;
}

int Engine::grow_desc(size_t n) {
    if (n <= desc_cap_) return GEPI_OK;
    const size_t c = std::max<size_t>(n, 2 * desc_cap_);
    drop_graphs(); /* they hold the old tables' addresses */
    for (int i = 0; i < NGROUP; i++) {
        hipFree(d_bdesc_[i]);
        hipFree(d_map_[i]);
        d_bdesc_[i] = nullptr;
        d_map_[i] = nullptr;
    }
    hipFree(d_edesc_);
    hipFree(d_tdesc_);
    hipFree(d_rep_);
    d_edesc_ = nullptr;
    d_tdesc_ = nullptr;
    d_rep_ = nullptr;
    desc_cap_ = 0;
    for (int i = 0; i < NGROUP; i++) {
        HIP_TRY(hipMalloc(&d_bdesc_[i], sizeof(BatchDesc) * c));
        HIP_TRY(hipMalloc(&d_map_[i], sizeof(int64_t) * 8 * (2 * c + 1)));
    }
    HIP_TRY(hipMalloc(&d_edesc_, sizeof(ExactDesc) * c));
    HIP_TRY(hipMalloc(&d_tdesc_, sizeof(TopkDesc) * c));
    HIP_TRY(hipMalloc(&d_rep_, sizeof(ItemReport) * c));
    desc_cap_ = c;
    return GEPI_OK;
}

size_t Engine::estimate_footprint(const gepi_panel *p, int64_t topmost) {
    const int64_t n_var_pad = ((p->n_var + TM - 1) / TM) * TM;
    /* super-steps of a scan whose classes hold every sample (a sample in
     * both masks, counted twice, is not assumed) */
    const int64_t n_u = ((p->n_samples + 127) / 128 + 2) / 3 + 2;
    const int64_t R = 2 * p->n_var;
    const int64_t k = std::min<int64_t>(topmost, 5 * (R * (R + 1) / 2));
    const unsigned long long cap =
        std::max<unsigned long long>((unsigned long long)k + TILE_EVALS_MAX + 1024, 4ull * k);
    const unsigned long long pcap = std::max<unsigned long long>(cap, 1ull << 20);
    return round_block((size_t)n_u * n_var_pad * 3 * 64) + round_block(cap * sizeof(Cand)) +
           round_block(pcap * sizeof(Pend)) + round_block((size_t)4 * p->w64 * sizeof(CompressWord)) +
           round_block(sizeof(ScanState));
}

void Engine::drop(ScanItem *it, int status) {
    dfree(it->d_layout);
    dfree(it->d_cw);
    dfree(it->d_pend);
    dfree(it->d_cand);
    dfree(it->d_state);
    it->d_layout = nullptr;
    it->d_cw = nullptr;
    it->d_pend = nullptr;
    it->d_cand = nullptr;
    it->d_state = nullptr;
    it->footprint = 0;
    it->status = status;
    it->done = true;
    it->stats.ms_scan = ms_since(it->t_added);
}

int Engine::add(ScanItem *it) {
    it->t_added = std::chrono::steady_clock::now();
    it->done = false;
    it->status = GEPI_OK;
    it->results.clear();
    memset(&it->stats, 0, sizeof it->stats);
    gepi_panel *p = it->panel;
    if (!p) {
        it->status = gepi_fail(GEPI_ERR_ARG, "null panel");
        it->done = true;
        return it->status;
    }
    if (p->device != device || it->topmost < 0 ||
        (it->ffun != GEPI_FFUNC_ENTROPY && it->ffun != GEPI_FFUNC_FISHERPV)) {
        it->status = gepi_fail(GEPI_ERR_ARG, "scan item: panel on device %d (engine %d), "
                                             "topmost %lld, objective %d",
                               p->device, device, (long long)it->topmost, it->ffun);
        it->done = true;
        return it->status;
    }
    build_compress_words(it->mask_ctrl, it->mask_case, p->n_samples, p->w64, it->h_cw, it->n0,
                         it->n1);
    it->n_var = p->n_var;
    const int64_t R = 2 * it->n_var;
    const int64_t total = 5 * (R * (R + 1) / 2);
    it->stats.evals = total;
    /* the job shape decides the class layout (wave tiles read the grouped
     * one): Fisher scans run wave tiles, which park flagged pairs; entropy
     * scans their class's tile (no table: batch_shape_default); captured
     * rounds have no wave-tile launch                                      */
    const int32_t ks = (it->n0 + 127) / 128 + (it->n1 + 127) / 128;
    const gepi_tune_class *tc = tune_of(ks);
    it->shape = tc ? tc->tile : batch_shape_default(ks);
    if (it->ffun != GEPI_FFUNC_ENTROPY && !use_graph_) it->shape = 2;
    if (it->shape == 2 && use_graph_) it->shape = 0;
    it->stats.count_path = it->shape == 2 ? 3 : 4;
    it->stats.n_class0 = it->n0;
    it->stats.n_class1 = it->n1;
    if ((int64_t)it->n0 + it->n1 > GEPI_MAX_SAMPLES) {
        it->status = gepi_fail(GEPI_ERR_LIMIT, "%d class samples exceed the limit of %d",
                               it->n0 + it->n1, GEPI_MAX_SAMPLES);
        it->done = true;
        return it->status;
    }
    it->k = std::min<int64_t>(it->topmost, total);
    if (it->k > (int64_t)(1u << 30)) {
        it->status = gepi_fail(GEPI_ERR_LIMIT, "topmost too large");
        it->done = true;
        return it->status;
    }
    if (it->k == 0) {
        it->done = true;
        it->stats.ms_scan = ms_since(it->t_added);
        return GEPI_OK;
    }
    int rc;
    if ((rc = ensure_tables(tabs_, it->n0 + it->n1 + 1, stream)) != GEPI_OK) {
        it->status = rc;
        it->done = true;
        return rc;
    }
    const int32_t w0 = ((it->n0 + 127) / 128) * 4, w1 = ((it->n1 + 127) / 128) * 4;
    it->n_k4 = (w0 + w1) / 4;
    it->ks0 = w0 / 4;
    it->n_var_pad = ((it->n_var + TM - 1) / TM) * TM;
    const int64_t n_a4 = it->n_var_pad / TM;
    it->n_tiles = n_a4 * (n_a4 + 1);
    const unsigned long long cap_min = (unsigned long long)it->k + TILE_EVALS_MAX + 1024;
    it->cap = std::max<unsigned long long>(cap_min, 4ull * (unsigned long long)it->k);
    if (getenv("GEPI_CAND_CAP")) it->cap = std::max(env_ull("GEPI_CAND_CAP", 0), cap_min);
    /* a pending list of at least 1 Mi survivors (16 MiB) lets loosely
     * filtered early bands grow without filling the candidate buffer      */
    it->pcap = std::max<unsigned long long>(it->cap, 1ull << 20);
    if (getenv("GEPI_PEND_CAP"))
        it->pcap = std::max<unsigned long long>(env_ull("GEPI_PEND_CAP", 0),
                                                (unsigned long long)TILE_EVALS_MAX + 1024);
    /* three K steps per super-step (f4_super3, see count_f4_run) */
    const int32_t n_u = f4_super3(w0 / 4) + f4_super3(w1 / 4); /* 64-byte rows each */
    const size_t lay = (size_t)std::max(n_u, 1) * it->n_var_pad * 3 * 4 * sizeof(uint4);
    it->d_layout = (uint4 *)dalloc(lay);
    it->d_cw = (CompressWord *)dalloc(it->h_cw.size() * sizeof(CompressWord));
    it->d_pend = (Pend *)dalloc(it->pcap * sizeof(Pend));
    it->d_cand = (Cand *)dalloc(it->cap * sizeof(Cand));
    it->d_state = (ScanState *)dalloc(sizeof(ScanState));
    if (!it->d_layout || !it->d_cw || !it->d_pend || !it->d_cand || !it->d_state) {
        drop(it, gepi_fail(GEPI_ERR_NOMEM, "device memory for a scan of %lld variants",
                           (long long)it->n_var));
        return it->status;
    }
    it->footprint = round_block(lay) + round_block(it->h_cw.size() * sizeof(CompressWord)) +
                    round_block(it->pcap * sizeof(Pend)) + round_block(it->cap * sizeof(Cand)) +
                    round_block(sizeof(ScanState));
    memset(&it->h_state0, 0, sizeof it->h_state0);
    it->h_state0.thr_score = INFINITY;
    it->h_state0.thr_idx = ~0ull;
    HIP_TRY(hipMemcpyAsync(it->d_cw, it->h_cw.data(), it->h_cw.size() * sizeof(CompressWord),
                           hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(it->d_state, &it->h_state0, sizeof(ScanState), hipMemcpyHostToDevice,
                           stream));
    if (it->n_k4 > 0) {
        const int64_t rows = it->n_var_pad * 3;
        const unsigned cgrid = (unsigned)((rows + 255) / 256);
        if (it->shape == 2)
            k_class_f4<3, true><<<cgrid, 256, 0, stream>>>(p->d_nat, (int32_t)it->n_var, p->w64,
                                                           it->d_cw, rows, f4_super3(it->ks0),
                                                           it->d_layout, nullptr);
        else
            k_class_f4<3><<<cgrid, 256, 0, stream>>>(p->d_nat, (int32_t)it->n_var, p->w64, it->d_cw,
                                                     rows, f4_super3(it->ks0), it->d_layout,
                                                     nullptr);
        HIP_TRY(hipGetLastError());
    }
    const double band_evals = tc ? tc->band_evals : 2.0e9;
    it->max_band = std::max<int64_t>(64, (int64_t)(band_evals / (double)F4_TILE_EVALS));
    it->band = std::max<int64_t>(1, std::min<int64_t>(it->max_band, (2 * it->k) / F4_TILE_EVALS + 1));
    it->forced_band = 0;
    if (const char *env = getenv("GEPI_BAND_TILES")) {
        it->forced_band = std::max<long long>(1, atoll(env));
        it->band = it->forced_band;
    }
    it->t0 = 0;
    it->count = 0;
    it->exact_total = 0;
    active_.push_back(it);
    return GEPI_OK;
}

int Engine::finish(ScanItem *it) {
    std::vector<Cand> host((size_t)it->count);
    if (it->count > 0)
        HIP_TRY(hipMemcpyAsync(host.data(), it->d_cand, sizeof(Cand) * it->count,
                               hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    it->results.resize((size_t)it->k);
    const int64_t n = candidates_to_results(host, it->k, it->results.data());
    it->results.resize((size_t)n);
    it->stats.exact_evals = (int64_t)it->exact_total;
    drop(it, n == it->k ? GEPI_OK
                        : gepi_fail(GEPI_ERR_INTERNAL, "top-k extraction returned %lld of %lld",
                                    (long long)n, (long long)it->k));
    return GEPI_OK;
}

int Engine::round(std::vector<ScanItem *> &finished) {
    if (active_.empty()) return GEPI_OK;
    const size_t na = active_.size();
    int rc;
    if ((rc = grow_desc(na)) != GEPI_OK) return rc;
    /* bands, scaled down uniformly when the round's workgroup budget is set */
    int64_t total_wg = 0;
    for (ScanItem *it : active_) {
        it->band_used = std::min<int64_t>(it->band, it->n_tiles - it->t0);
        total_wg += it->band_used;
    }
    if (round_wg_budget > 0 && total_wg > round_wg_budget) {
        const double f = (double)round_wg_budget / (double)total_wg;
        for (ScanItem *it : active_)
            it->band_used = std::max<int64_t>(1, (int64_t)((double)it->band_used * f));
    }
    /* count launch groups: objective (g & 1) x job shape (g >> 1) */
    std::vector<BatchDesc> bd[NGROUP];
    std::vector<ScanItem *> grp[NGROUP];
    int64_t wg[NGROUP] = {};
    std::vector<ExactDesc> ed(na);
    std::vector<TopkDesc> td(na);
    for (size_t i = 0; i < na; i++) {
        ScanItem *it = active_[i];
        grp[(it->ffun == GEPI_FFUNC_ENTROPY ? 0 : 1) + 2 * it->shape].push_back(it);
        ExactDesc &e = ed[i];
        memset(&e, 0, sizeof e);
        e.e.st = it->d_state;
        e.e.pend = it->d_pend;
        e.e.pcap = it->pcap;
        e.e.cand = it->d_cand;
        e.e.cap = it->cap;
        e.e.lhi = tabs_.lhi;
        e.e.llo = tabs_.llo;
        e.ffun = it->ffun;
        TopkDesc &t = td[i];
        t.st = it->d_state;
        t.cand = it->d_cand;
        t.cap = it->cap;
        t.drop_t0 = it->t0;
        t.drop_t1 = it->t0 + it->band_used;
        t.report = d_rep_ + i;
        t.k = (uint32_t)it->k;
        t.ffun = it->ffun;
        t.n_a4 = (int32_t)(it->n_var_pad / TM);
        t.pad = 0;
    }
    /* per objective: descriptors with the longest K loops first (their
     * workgroups are dispatched first; short ones fill the tail), and the
     * per-XCD chunk tables of BatchMap */
    /* launch form: the one of the class with the most count work (tiles x
     * K steps) in this round                                               */
    int per_cu_dflt = 2;
    if (!tuning.empty()) {
        std::vector<double> work(tuning.size(), 0.0);
        for (ScanItem *it : active_)
            work[(size_t)(tune_of(it->n_k4) - tuning.data())] +=
                (double)it->band_used * (double)std::max(it->n_k4, 1);
        const size_t c = (size_t)(std::max_element(work.begin(), work.end()) - work.begin());
        per_cu_dflt = tuning[c].wg_per_cu;
    }
    const int per_cu = f4_wg_per_cu(per_cu_dflt);
    /* captured rounds need a fixed grid: the persistent launch form */
    const bool graph = use_graph_ && per_cu > 0;
    /* Fisher scans on wave tiles park flagged pairs in regions of one pool */
    uint32_t *park_busy = nullptr;
    if (!grp[5].empty()) {
        const int32_t slots = (int32_t)((4 * (int64_t)device_cus(device) + 7) / 8 * 8);
        const size_t rec_bytes = (size_t)slots * F4W_PAIR_CAP * sizeof(PairRec);
        if (park_slots_ != slots) {
            hipFree(d_park_);
            d_park_ = nullptr;
            park_slots_ = 0;
            HIP_TRY(hipMalloc(&d_park_, rec_bytes + (size_t)(slots + 8) * sizeof(uint32_t)));
            park_slots_ = slots;
        }
        park_busy = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(d_park_) + rec_bytes);
        HIP_TRY(hipMemsetAsync(park_busy, 0, (size_t)(slots + 8) * sizeof(uint32_t), stream));
    }
    std::vector<int64_t> map[NGROUP];
    for (int f = 0; f < NGROUP; f++) {
        std::vector<ScanItem *> &g = grp[f];
        /* jobs: tiles (shape 0), half tiles (1: two per tile, numbered
         * 2 * tile + half) or units of two tiles (2: the descriptor still
         * counts tiles)                                                     */
        const int shape = f >> 1;
        const int64_t per_tile = shape == 1 ? 2 : 1;
        std::stable_sort(g.begin(), g.end(),
                         [](const ScanItem *a, const ScanItem *b) { return a->n_k4 > b->n_k4; });
        const size_t n = g.size();
        const size_t ns = graph ? desc_cap_ : n; /* scan slots of the map */
        map[f].assign(8 * (ns + 1) + 8 * ns, 0);
        int64_t *xpre = map[f].data(), *xst = map[f].data() + 8 * (ns + 1);
        int64_t dealt = 0;
        for (size_t j = 0; j < n; j++) {
            ScanItem *it = g[j];
            BatchDesc d;
            memset(&d, 0, sizeof d);
            d.g.planes_k = reinterpret_cast<const uint64_t *>(it->d_layout);
            d.g.pc3 = nullptr; /* three-K-step layout: no correction */
            d.g.tile_begin = it->t0 * per_tile;
            d.g.n_tiles = (int32_t)(it->band_used * per_tile);
            d.g.n_var = (int32_t)it->n_var;
            d.g.n_var_pad = (int32_t)it->n_var_pad;
            d.g.n_ks = it->n_k4;
            d.g.ks0 = it->ks0;
            d.g.f32 = f32_counts(it->n0, it->n1) ? 1 : 0;
            d.g.n0 = it->n0;
            d.g.n1 = it->n1;
            d.g.st = it->d_state;
            d.g.pend = it->d_pend;
            d.g.pcap = it->pcap;
            d.g.lfact = tabs_.lfact;
            if (f == 5) {
                d.g.park = d_park_;
                d.g.park_busy = park_busy;
                d.g.park_cap = F4W_PAIR_CAP;
                d.g.park_slots = (uint32_t)park_slots_;
            }
            bd[f].push_back(d);
            const int64_t b = shape == 2 ? (it->band_used + 1) / 2 : it->band_used * per_tile;
            const int64_t r = dealt & 7;
            int64_t start = 0;
            for (int x = 0; x < 8; x++) {
                const int64_t chunk = b / 8 + (((x - r) & 7) < (b & 7) ? 1 : 0);
                xst[(size_t)x * ns + j] = start;
                xpre[(size_t)x * (ns + 1) + j + 1] = xpre[(size_t)x * (ns + 1) + j] + chunk;
                start += chunk;
            }
            dealt += b;
        }
        /* slots past the round's scans: empty ranges */
        for (int x = 0; x < 8; x++)
            for (size_t j = n; j < ns; j++)
                xpre[(size_t)x * (ns + 1) + j + 1] = xpre[(size_t)x * (ns + 1) + n];
        wg[f] = dealt;
    }
    if (graph) {
        /* every descriptor slot is launched: the ones past the round's scans
         * are zeroed (no scan)                                              */
        ed.resize(desc_cap_);
        td.resize(desc_cap_);
    }
    for (int f = 0; f < NGROUP; f++) {
        if (!bd[f].empty())
            HIP_TRY(hipMemcpyAsync(d_bdesc_[f], bd[f].data(), sizeof(BatchDesc) * bd[f].size(),
                                   hipMemcpyHostToDevice, stream));
        if (!bd[f].empty() || graph)
            HIP_TRY(hipMemcpyAsync(d_map_[f], map[f].data(), sizeof(int64_t) * map[f].size(),
                                   hipMemcpyHostToDevice, stream));
    }
    HIP_TRY(hipMemcpyAsync(d_edesc_, ed.data(), sizeof(ExactDesc) * ed.size(), hipMemcpyHostToDevice,
                           stream));
    HIP_TRY(hipMemcpyAsync(d_tdesc_, td.data(), sizeof(TopkDesc) * td.size(), hipMemcpyHostToDevice,
                           stream));
    for (int f = 0; f < NGROUP; f++) wq_parity_[f] = 0;
    if (graph) {
        if (gx_cap_ != desc_cap_ || gx_per_cu_ != per_cu) {
            drop_graphs();
            gx_cap_ = desc_cap_;
            gx_per_cu_ = per_cu;
        }
        int mask = 0;
        for (int f = 0; f < 4; f++) mask |= wg[f] > 0 ? 1 << f : 0; /* no wave tiles here */
        if (!gx_count_[mask] && (rc = capture_graph(mask, per_cu)) != GEPI_OK) return rc;
        if (!gx_finish_ && (rc = capture_graph(NGRAPH_COUNT, per_cu)) != GEPI_OK) return rc;
        /* the count graph clears the work counters first */
        HIP_TRY(hipEventRecord(ev0_, stream));
        HIP_TRY(hipGraphLaunch(gx_count_[mask], stream));
        HIP_TRY(hipEventRecord(ev1_, stream));
        HIP_TRY(hipGraphLaunch(gx_finish_, stream));
        graph_rounds++;
    } else {
        /* work counters start clean every round (a failed round leaves no
         * stale counts behind) */
        HIP_TRY(hipMemsetAsync(d_wq_, 0, sizeof(uint32_t) * 16 * NGROUP, stream));
        HIP_TRY(hipEventRecord(ev0_, stream));
        for (int f = 0; f < NGROUP; f++) {
            if (wg[f] == 0) continue;
            const size_t n = bd[f].size();
            BatchMap m;
            m.xprefix = d_map_[f];
            m.xstart = d_map_[f] + 8 * (n + 1);
            m.n = (int32_t)n;
            m.wq = d_wq_ + 16 * f;
            m.wq_parity = wq_parity_[f];
            /* one slot per workgroup: the grid covers the largest XCD share */
            int64_t xmax = 0;
            for (int x = 0; x < 8; x++) xmax = std::max<int64_t>(xmax, map[f][(size_t)x * (n + 1) + n]);
            unsigned grid = (unsigned)(8 * xmax);
            if (per_cu > 0 && f < 4) {
                wq_parity_[f] ^= 1;
                grid = (unsigned)std::min<int64_t>(
                    std::max<int64_t>(8, (per_cu * (int64_t)device_cus(device)) / 8 * 8),
                    ((wg[f] + 7) / 8) * 8);
            }
            launch_count(f, per_cu > 0, grid, m);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ev1_, stream));
        k_exact_batch<<<dim3(EXACT_BLOCKS_X, (unsigned)na), 256, 0, stream>>>(d_edesc_);
        k_topk_wg<<<(unsigned)na, TOPK_THREADS, 0, stream>>>(d_tdesc_);
        HIP_TRY(hipGetLastError());
    }
    std::vector<ItemReport> rep(na);
    HIP_TRY(hipMemcpyAsync(rep.data(), d_rep_, sizeof(ItemReport) * na, hipMemcpyDeviceToHost,
                           stream));
    HIP_TRY(hipStreamSynchronize(stream));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev0_, ev1_) != hipSuccess) {
        /* a statistic only: leave no error behind for the next call */
        ms = 0.0f;
        (void)hipGetLastError();
    }
    ms_count += ms;
    rounds++;

    std::vector<ScanItem *> still;
    still.reserve(na);
    for (size_t i = 0; i < na; i++) {
        ScanItem *it = active_[i];
        const ItemReport &r = rep[i];
        const int64_t used = it->band_used;
        it->stats.bands++;
        if (r.pend_overflow) {
            it->stats.pend_overflows++;
            it->band = std::max<int64_t>(1, used / 4);
            still.push_back(it);
            continue;
        }
        if (r.cand_overflow) {
            it->stats.cand_overflows++;
            it->count = r.count;
            it->band = std::max<int64_t>(1, used / 4);
            still.push_back(it);
            continue;
        }
        it->exact_total += r.pend_count;
        const unsigned long long added = r.raw_count - it->count;
        it->count = r.count;
        /* next band: survivors filling at most half the pending list and
         * admissions at most half the candidate room, at this band's rates */
        const double pend_rate = (double)r.pend_count / (double)used;
        const double adm_rate = (double)added / (double)used;
        int64_t next = it->max_band;
        if (pend_rate > 0.0) next = std::min<int64_t>(next, (int64_t)((double)(it->pcap / 2) / pend_rate));
        if (adm_rate > 0.0)
            next = std::min<int64_t>(
                next, (int64_t)((double)((it->cap - (unsigned long long)it->k) / 2) / adm_rate));
        it->band = std::max<int64_t>(1, std::min<int64_t>(next, used * 4));
        if (it->forced_band) it->band = std::min<int64_t>(it->forced_band, used * 4);
        it->t0 += used;
        if (it->t0 >= it->n_tiles) {
            if ((rc = finish(it)) != GEPI_OK) return rc;
            finished.push_back(it);
        } else {
            still.push_back(it);
        }
    }
    active_.swap(still);
    return GEPI_OK;
}

} // namespace gepi_host

using namespace gepi_host;

struct gepi_engine {
    Engine eng;
};

extern "C" int gepi_engine_create(int32_t device, gepi_engine **out) {
    if (!out) return gepi_fail(GEPI_ERR_ARG, "null out");
    *out = nullptr;
    gepi_engine *e = new gepi_engine();
    int rc = e->eng.init(device);
    if (rc != GEPI_OK) {
        e->eng.release();
        delete e;
        return rc;
    }
    *out = e;
    return GEPI_OK;
}

extern "C" int gepi_engine_destroy(gepi_engine *e) {
    if (!e) return GEPI_OK;
    e->eng.release();
    delete e;
    return GEPI_OK;
}

extern "C" int gepi_engine_autotune(gepi_engine *e, int32_t repeats, gepi_tune_class *out,
                                    int32_t n_out, int32_t *n_classes) {
    if (!e) return gepi_fail(GEPI_ERR_ARG, "null engine");
    if (n_out < 0 || (n_out > 0 && !out)) return gepi_fail(GEPI_ERR_ARG, "bad output table");
    if (!e->eng.idle()) return gepi_fail(GEPI_ERR_ARG, "engine busy");
    int rc;
    if ((rc = set_device(e->eng.device)) != GEPI_OK) return rc;
    if ((rc = e->eng.autotune(repeats)) != GEPI_OK) return rc;
    const int32_t n = (int32_t)e->eng.tuning.size();
    for (int32_t i = 0; i < std::min(n, n_out); i++) out[i] = e->eng.tuning[(size_t)i];
    if (n_classes) *n_classes = n;
    return GEPI_OK;
}

extern "C" int gepi_engine_set_tuning(gepi_engine *e, const gepi_tune_class *classes, int32_t n) {
    if (!e) return gepi_fail(GEPI_ERR_ARG, "null engine");
    if (n < 0 || n > 16 || (n > 0 && !classes)) return gepi_fail(GEPI_ERR_ARG, "bad class count %d", n);
    for (int32_t i = 0; i < n; i++) {
        const gepi_tune_class &c = classes[i];
        if ((i > 0 && c.max_ksteps <= classes[i - 1].max_ksteps) || c.max_ksteps < 1)
            return gepi_fail(GEPI_ERR_ARG, "class %d: max_ksteps must increase", i);
        if (c.wg_per_cu < 0 || c.wg_per_cu > 16)
            return gepi_fail(GEPI_ERR_ARG, "class %d: wg_per_cu %d outside [0, 16]", i, c.wg_per_cu);
        if (c.tile < 0 || c.tile > 2)
            return gepi_fail(GEPI_ERR_ARG, "class %d: tile %d is not 0, 1 or 2", i, c.tile);
        if (!(c.band_evals >= 1e6) || !std::isfinite(c.band_evals))
            return gepi_fail(GEPI_ERR_ARG, "class %d: band_evals below 1e6", i);
    }
    if (!e->eng.idle()) return gepi_fail(GEPI_ERR_ARG, "engine busy");
    e->eng.tuning.assign(classes, classes + n);
    return GEPI_OK;
}

extern "C" int gepi_engine_scan_batch(gepi_engine *e, gepi_batch_item *items, int32_t n_items,
                                      gepi_batch_stats *stats) {
    if (!e) return gepi_fail(GEPI_ERR_ARG, "null engine");
    if (n_items < 0 || (n_items > 0 && !items)) return gepi_fail(GEPI_ERR_ARG, "bad item list");
    int rc;
    if ((rc = set_device(e->eng.device)) != GEPI_OK) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    gepi_batch_stats bs;
    memset(&bs, 0, sizeof bs);
    std::vector<ScanItem> its((size_t)n_items);
    e->eng.ms_count = 0.0;
    e->eng.rounds = 0;
    e->eng.graph_rounds = 0;
    for (int32_t i = 0; i < n_items; i++) {
        ScanItem &it = its[(size_t)i];
        it.panel = items[i].panel;
        it.mask_ctrl = items[i].mask_ctrl;
        it.mask_case = items[i].mask_case;
        it.topmost = items[i].topmost;
        it.ffun = items[i].ffun;
        items[i].n_tops = 0;
        if (it.topmost > 0 && !items[i].tops) {
            it.status = gepi_fail(GEPI_ERR_ARG, "null tops");
            it.done = true;
            continue;
        }
        e->eng.add(&it);
    }
    HIP_TRY(hipStreamSynchronize(e->eng.stream));
    bs.ms_prepare = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                        .count();
    std::vector<ScanItem *> fin;
    rc = GEPI_OK;
    while (!e->eng.idle()) {
        if ((rc = e->eng.round(fin)) != GEPI_OK) break;
    }
    for (int32_t i = 0; i < n_items; i++) {
        ScanItem &it = its[(size_t)i];
        if (!it.done) it.status = rc != GEPI_OK ? rc : GEPI_ERR_INTERNAL;
        items[i].status = it.status;
        items[i].stats = it.stats;
        if (it.status == GEPI_OK) {
            const int64_t n = (int64_t)it.results.size();
            if (n > 0) memcpy(items[i].tops, it.results.data(), sizeof(gepi_result) * (size_t)n);
            items[i].n_tops = n;
        } else {
            bs.items_failed++;
        }
        bs.evals += it.stats.evals;
    }
    bs.ms_count_kernels = e->eng.ms_count;
    bs.rounds = e->eng.rounds;
    bs.graph_rounds = e->eng.graph_rounds;
    bs.ms_batch = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                      .count();
    if (stats) *stats = bs;
    return rc;
}
