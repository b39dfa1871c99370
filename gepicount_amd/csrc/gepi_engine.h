/*
 * gepi_engine.h -- the batched scan engine (host side).  Scans are added at
 * any time and advance together in rounds; a round is one count launch per
 * objective and job shape present, one exact-score launch and one top-k
 * launch for all active
 * scans, then a single read of their reports; with GEPI_ENGINE_GRAPH=1 a
 * round's launches are replayed from two captured HIP graphs.  The batch C
 * ABI runs rounds
 * until every scan is done; the serving loop admits new requests between
 * rounds.  All device work runs on the engine's stream.
 */
#ifndef GEPI_ENGINE_H
#define GEPI_ENGINE_H

#include <hip/hip_runtime.h>

#include <chrono>
#include <map>
#include <vector>

#include "../../include/gepicount.h"
#include "gepi_host.h"

namespace gepi {
struct BatchDesc;
struct BatchMap;
struct ExactDesc;
struct TopkDesc;
struct ItemReport;
struct PairRec;
} // namespace gepi

namespace gepi_host {

/* One scan inside the engine. */
struct ScanItem {
    /* request (masks are read only by Engine::add) */
    gepi_panel *panel = nullptr;
    const uint64_t *mask_ctrl = nullptr, *mask_case = nullptr;
    int64_t topmost = 0;
    int32_t ffun = GEPI_FFUNC_ENTROPY;
    void *tag = nullptr; /* caller's */
    /* outcome */
    int32_t status = GEPI_OK;
    bool done = false;
    std::vector<gepi_result> results;
    gepi_scan_stats stats{};
    std::chrono::steady_clock::time_point t_added;
    /* scan state */
    int64_t k = 0, n_var = 0, n_var_pad = 0, n_tiles = 0, t0 = 0;
    int64_t band = 1, band_used = 0, max_band = 1, forced_band = 0;
    int32_t n0 = 0, n1 = 0, n_k4 = 0, ks0 = 0;
    /* count job shape (its class's tile): 0 tiles, 1 half tiles, 2 wave
     * tiles on the grouped layout                                          */
    int32_t shape = 0;
    unsigned long long cap = 0, pcap = 0, count = 0, exact_total = 0;
    uint4 *d_layout = nullptr;
    gepi::CompressWord *d_cw = nullptr;
    gepi::Pend *d_pend = nullptr;
    gepi::Cand *d_cand = nullptr;
    gepi::ScanState *d_state = nullptr;
    size_t footprint = 0; /* device bytes held */
    std::vector<gepi::CompressWord> h_cw;
    gepi::ScanState h_state0;
};

class Engine {
  public:
    int init(int device);
    void release();
    /* Validates and stages a scan (class layout launched on the stream).  A
     * scan that fails validation gets its status set and done = true.       */
    int add(ScanItem *it);
    /* One round over the active scans; completed scans (results filled,
     * device memory returned) are appended to `finished`.                    */
    int round(std::vector<ScanItem *> &finished);
    bool idle() const { return active_.empty(); }
    size_t active_count() const { return active_.size(); }
    /* Device bytes a scan of this panel would hold (admission control). */
    static size_t estimate_footprint(const gepi_panel *p, int64_t topmost);
    size_t bytes_in_use() const { return in_use_; }
    /* Workgroups per count round (0 = unlimited): bounds round latency. */
    int64_t round_wg_budget = 0;
    double ms_count = 0.0;
    int32_t rounds = 0;
    int32_t graph_rounds = 0; /* rounds replayed from the captured graphs     */
    int device = 0;
    hipStream_t stream = nullptr;
    /* Schedule per size class, ascending max_ksteps (empty: 2 workgroups
     * per CU and 2e9 evaluations per band for every scan).                  */
    std::vector<gepi_tune_class> tuning;
    /* Measures the table on synthetic batches (gepi_engine_autotune). */
    int autotune(int32_t repeats);
    /* The last autotune's check on all trial scans together (ms; infinite
     * when the per-class table equalled the default).                     */
    double mixed_ms_tuned_ = 0.0, mixed_ms_default_ = 0.0;

  private:
    void *dalloc(size_t bytes);
    void dfree(void *p);
    int finish(ScanItem *it);
    void drop(ScanItem *it, int status);
    int grow_desc(size_t n);
    /* The class of a scan with n_k4 K steps (nullptr without a table). */
    const gepi_tune_class *tune_of(int32_t n_k4) const;
    /* Wall time (ms) of one synthetic batch: every item added, rounds
     * until idle.                                                           */
    int time_batch(std::vector<ScanItem> &items, double *ms);
    /* Captures part of a round for descriptor capacity desc_cap_ and a
     * persistent grid of per_cu workgroups per CU: part 0-15 into
     * gx_count_[part] (work-counter clear, then the count launches of the
     * groups in bit mask part: bit g = objective (g & 1: 0 entropy, 1
     * Fisher) and job shape (g >> 1: 0 whole tiles, 1 half tiles)), part
     * NGRAPH_COUNT into gx_finish_ (exact + top-k launches).  Wave-tile
     * groups (4, 5) have no persistent form and are never captured: with
     * graphs on, scans keep whole tiles instead.                          */
    int capture_graph(int part, int per_cu);
    /* One count launch of group g (as in capture_graph). */
    void launch_count(int g, bool persist, unsigned grid, const gepi::BatchMap &m);
    void drop_graphs();

    DeviceTables tabs_;
    std::vector<ScanItem *> active_;
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t in_use_ = 0;
    /* count launch groups: objective (g & 1) x job shape (g >> 1) */
    static constexpr int NGROUP = 6, NGRAPH_COUNT = 16;
    gepi::BatchDesc *d_bdesc_[NGROUP] = {};
    int64_t *d_map_[NGROUP] = {};
    gepi::ExactDesc *d_edesc_ = nullptr;
    gepi::TopkDesc *d_tdesc_ = nullptr;
    gepi::ItemReport *d_rep_ = nullptr;
    size_t desc_cap_ = 0;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    uint32_t *d_wq_ = nullptr; /* fp4 work counters: [group][2 sets][8] */
    int32_t wq_parity_[NGROUP] = {};
    /* Fisher's parked variant pairs on wave tiles: park_slots_ regions of
     * F4W_PAIR_CAP records, then park_slots_ + 8 busy / rotation words     */
    gepi::PairRec *d_park_ = nullptr;
    int32_t park_slots_ = 0;
    /* rounds as captured graphs (GEPI_ENGINE_GRAPH=1; else launched one by
     * one); valid for gx_cap_ descriptor slots and gx_per_cu_ workgroups
     * per CU                                                                 */
    bool use_graph_ = false;
    hipGraphExec_t gx_count_[NGRAPH_COUNT] = {}, gx_finish_ = nullptr;
    size_t gx_cap_ = 0;
    int gx_per_cu_ = 0;
};

} // namespace gepi_host

#endif
