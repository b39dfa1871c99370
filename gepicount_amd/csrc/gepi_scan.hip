/*
 * gepi_scan.hip -- host orchestration and C ABI of libgepicount.
 *
 * Replaces launch_process_gpu_1 and process_gpu_1 (src/kernels/kernel1.cu:295-516),
 * the result-buffer management of src/kernels/misc1.cu:35-295 and the panel
 * upload of src/kernels/setup1.h.  Kernels live in gepi_stage.h (panel
 * staging), gepi_count.h (count + filter) and gepi_topk.h (exact scores and
 * streaming top-k).
 *
 * Data layout in HBM
 *   panel    [n_var][3][W64] u64: per variant the allele-1 plane (reader row
 *            2i masked by valid), allele-2 plane (row 2i+1) and valid plane,
 *            file sample order.  Resident across scans.
 *   classes  [n_var_pad][3][wc] u32 per scan: samples reordered so that class
 *            0 (controls) then class 1 (cases) are contiguous, each segment
 *            zero-padded to 128 bits; class masks become word ranges.
 *   classes_k [n_ks][n_var_pad][3] u64 (MFMA path): the same bits K-major.
 *   pend     [cap] 16-byte filter survivors; cand [2][cap] 32-byte candidates.
 *
 * A scan walks the lower triangle of variant tiles in bands: count kernel
 * (filter survivors -> pend), k_exact (exact score, admit keys <= threshold
 * -> cand), then, when more than k candidates exist, a radix select of the
 * k-th key, compaction to exactly k records and a new threshold.  The host
 * synchronises once per band to read the state and size the next band; only
 * the final k records are copied to the host.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gepicount.h"
#include "gepi_count.h"
#include "gepi_internal.h"
#include "gepi_score.h"
#include "gepi_stage.h"
#include "gepi_topk.h"
#include "gepi_host.h"

#define GEPI_VERSION "gepicount-mi355x 0.2.0"

using namespace gepi;
using namespace gepi_host;

/* ------------------------------------------------------------------------ */
/* errors                                                                    */
static thread_local char g_err[1024] = "";

int gepi_fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *gepi_last_error(void) { return g_err; }
extern "C" const char *gepi_version(void) { return GEPI_VERSION; }

/* ------------------------------------------------------------------------ */
/* device, tables                                                            */

namespace gepi_host {

/* compute units per device, recorded by set_device */
static int32_t g_cus[64];

int32_t device_cus(int device) {
    return device >= 0 && device < 64 && g_cus[device] > 0 ? g_cus[device] : 256;
}

int f4_wg_per_cu(int dflt) {
    /* launch form of the batched fp4 count kernel: GEPI_F4_WG_PER_CU = k > 0
     * runs k persistent workgroups per CU, 0 one tile per workgroup */
    const char *v = getenv("GEPI_F4_WG_PER_CU");
    if (!v) return dflt;
    const int k = atoi(v);
    return k < 0 ? 0 : k > 16 ? 16 : k;
}

int set_device(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return gepi_fail(GEPI_ERR_DEVICE, "no HIP device available (%s)",
                         e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    if (device < 0 || device >= n)
        return gepi_fail(GEPI_ERR_ARG, "device %d out of range (%d visible)", device, n);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return gepi_fail(GEPI_ERR_DEVICE, "device %d is %s; this build targets gfx950", device,
                         prop.gcnArchName);
    if (device < 64) g_cus[device] = prop.multiProcessorCount;
    return GEPI_OK;
}

int ensure_tables(DeviceTables &tabs, int32_t tmax, hipStream_t stream) {
    if (tmax <= tabs.tmax) return GEPI_OK;
    int32_t n = tmax + 1;
    std::vector<double> hi(n), lo(n), lf(n);
    gepi_build_tables(n, hi.data(), lo.data(), lf.data());
    /* the old tables may still be read by work queued on the stream */
    HIP_TRY(hipStreamSynchronize(stream));
    hipFree(tabs.lhi);
    hipFree(tabs.llo);
    hipFree(tabs.lfact);
    tabs = DeviceTables();
    HIP_TRY(hipMalloc(&tabs.lhi, sizeof(double) * n));
    HIP_TRY(hipMalloc(&tabs.llo, sizeof(double) * n));
    HIP_TRY(hipMalloc(&tabs.lfact, sizeof(double) * n));
    /* copies on the stream that will read the tables (a null-stream copy is
     * not ordered with a non-blocking stream); the host vectors live until
     * the synchronisation below */
    HIP_TRY(hipMemcpyAsync(tabs.lhi, hi.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(tabs.llo, lo.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(tabs.lfact, lf.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                           stream));
    HIP_TRY(hipStreamSynchronize(stream));
    tabs.tmax = tmax;
    return GEPI_OK;
}

} // namespace gepi_host

/* The fixed scratch array of gepi_scan_noalloc (layout at bind_scratch).  */
constexpr size_t SCRATCH_BYTES = (size_t)GEPI_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

/* ------------------------------------------------------------------------ */
/* panel                                                                     */

namespace {

int panel_alloc(gepi_panel *p, int device, int64_t n_var, int32_t n_samples) {
    p->device = device;
    p->n_var = n_var;
    p->n_samples = n_samples;
    p->w64 = (n_samples + 63) / 64;
    size_t bytes = (size_t)(n_var > 0 ? n_var : 1) * 3 * (size_t)(p->w64 > 0 ? p->w64 : 1) * 8;
    HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&p->d_nat, bytes));
    HIP_TRY(hipMemsetAsync(p->d_nat, 0, bytes, p->stream));
    HIP_TRY(hipMalloc(&p->d_state, sizeof(ScanState)));
    HIP_TRY(hipMalloc(&p->d_small, sizeof(uint64_t) * 2 * SEL_SMALL));
    HIP_TRY(hipEventCreate(&p->ev0));
    HIP_TRY(hipEventCreate(&p->ev1));
    return GEPI_OK;
}

void panel_free(gepi_panel *p) {
    if (!p) return;
    hipSetDevice(p->device);
    if (p->stream) hipStreamSynchronize(p->stream);
    hipFree(p->d_nat);
    hipFree(p->d_cls);
    hipFree(p->d_clsk);
    hipFree(p->d_perm);
    hipFree(p->d_cw);
    hipFree(p->d_park);
    hipFree(p->d_cand[0]);
    hipFree(p->d_cand[1]);
    hipFree(p->d_pend);
    hipFree(p->d_state);
    hipFree(p->d_small);
    hipFree(p->tabs.lhi);
    hipFree(p->tabs.llo);
    hipFree(p->tabs.lfact);
    if (p->ev0) hipEventDestroy(p->ev0);
    if (p->ev1) hipEventDestroy(p->ev1);
    if (p->stream) hipStreamDestroy(p->stream);
    delete p;
}

int check_pairing(const uint64_t *valid, int64_t n_rows, int32_t words) {
    for (int64_t i = 0; i + 1 < n_rows; i += 2) {
        const uint64_t *a = valid + (size_t)i * words, *b = a + words;
        if (memcmp(a, b, sizeof(uint64_t) * (size_t)words) != 0)
            return gepi_fail(GEPI_ERR_LAYOUT,
                             "rows %lld and %lld do not share valid bits; rows must come in "
                             "allele pairs of one variant",
                             (long long)i, (long long)(i + 1));
    }
    return GEPI_OK;
}

int validate_params(const gepi_params *par) {
    if (!par) return gepi_fail(GEPI_ERR_ARG, "null params");
    if (par->topmost < 0) return gepi_fail(GEPI_ERR_ARG, "topmost must be >= 0");
    if (par->ffun != GEPI_FFUNC_ENTROPY && par->ffun != GEPI_FFUNC_FISHERPV)
        return gepi_fail(GEPI_ERR_ARG, "unknown objective %d", par->ffun);
    if (par->count_path < 0 || par->count_path > 4)
        return gepi_fail(GEPI_ERR_ARG, "unknown count path %d", par->count_path);
    return GEPI_OK;
}

/* Select the rank-th smallest key (1-based) among the count records of src. */
/* Device buffers of one scan: the panel's own (grown on demand) or carved
 * from the fixed scratch array of gepi_scan_noalloc. */
struct ScanBufs {
    uint32_t *cls = nullptr;
    uint64_t *clsk = nullptr;
    int32_t *perm = nullptr;
    CompressWord *cw = nullptr;
    Cand *cand[2] = {nullptr, nullptr};
    Pend *pend = nullptr;
    ScanState *state = nullptr;
    uint64_t *small = nullptr;
    const double *lhi = nullptr, *llo = nullptr, *lfact = nullptr;
};

void launch_select(gepi_panel *p, const ScanBufs &b, const Cand *src, uint32_t rank,
                   unsigned long long count) {
    k_sel_reset<<<1, 1, 0, p->stream>>>(b.state, rank);
    unsigned long long blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    for (int pass = 0; pass < 16; pass++) {
        k_sel_hist<<<(unsigned)blocks, 256, 0, p->stream>>>(src, b.state);
        k_sel_pick<<<1, 1, 0, p->stream>>>(b.state);
    }
    k_sel_gather<<<(unsigned)blocks, 256, 0, p->stream>>>(src, b.state, b.small);
    k_sel_small<<<1, SEL_SMALL, 0, p->stream>>>(b.state, b.small);
}

int read_state(gepi_panel *p, const ScanBufs &b, ScanState *s) {
    HIP_TRY(hipMemcpyAsync(s, b.state, sizeof(ScanState), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return GEPI_OK;
}

/* ---- fixed scratch of gepi_scan_noalloc ---------------------------------
 * `scratch` (file scope above): one array per GPU, held until the process
 * ends.  Layout: the score tables
 * (t log2 t high and low parts, ln t!) for t <= GEPI_MAX_SAMPLES + 1 at fixed
 * offsets, the scan state, the select buffer, then the per-call regions
 * (compression words, class layout, pending list, two candidate lists). In
 * host code `scratch` is only a symbol: its device address comes from
 * hipGetSymbolAddress, once per device.                                     */
constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }
constexpr int32_t SCR_TMAX = GEPI_MAX_SAMPLES + 1;
constexpr size_t SCR_TAB = align256(sizeof(double) * (size_t)(SCR_TMAX + 1));
constexpr size_t SCR_STATE = 3 * SCR_TAB;
constexpr size_t SCR_SMALL = SCR_STATE + align256(sizeof(ScanState));
constexpr size_t SCR_DYN = SCR_SMALL + align256(sizeof(uint64_t) * 2 * SEL_SMALL);
static_assert(SCR_DYN < SCRATCH_BYTES / 8, "fixed scratch part too large");

struct ScratchDevice {
    unsigned char *ptr = nullptr; /* device address of `scratch` on this GPU */
    int32_t tab_tmax = -1;        /* tables valid for t <= tab_tmax           */
};
ScratchDevice g_scratch[64];
std::mutex g_scratch_mu;

/* Binds b to the scratch array of p's device (current): tables grown to
 * tmax, then regions for the compression words, the class layout and the
 * lists.  Allocates nothing; fails with GEPI_ERR_LIMIT when they do not fit. */
int bind_scratch(gepi_panel *p, ScanBufs &b, int32_t tmax, size_t cw_bytes, size_t clsk_bytes,
                 unsigned long long cap, unsigned long long pcap) {
    const int dev = p->device;
    if (dev < 0 || dev >= 64) return gepi_fail(GEPI_ERR_DEVICE, "device %d out of range", dev);
    const size_t need = SCR_DYN + align256(cw_bytes) + align256(clsk_bytes) +
                        align256(sizeof(Pend) * pcap) + 2 * align256(sizeof(Cand) * cap);
    if (need > SCRATCH_BYTES)
        return gepi_fail(GEPI_ERR_LIMIT,
                         "gepi_scan_noalloc: this scan needs %zu bytes of scratch, more than its "
                         "%zu (class layout %zu, candidate lists %llu records); use gepi_scan",
                         need, SCRATCH_BYTES, clsk_bytes, cap);
    if (tmax > SCR_TMAX)
        return gepi_fail(GEPI_ERR_LIMIT, "gepi_scan_noalloc: %d class samples exceed %d", tmax - 1,
                         SCR_TMAX - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ScratchDevice &sd = g_scratch[dev];
    if (!sd.ptr) {
        void *q = nullptr;
        hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q)
            return gepi_fail(GEPI_ERR_DEVICE, "scratch address on device %d: %s", dev,
                             hipGetErrorString(e));
        sd.ptr = static_cast<unsigned char *>(q);
    }
    unsigned char *base = sd.ptr;
    double *lhi = reinterpret_cast<double *>(base);
    double *llo = reinterpret_cast<double *>(base + SCR_TAB);
    double *lf = reinterpret_cast<double *>(base + 2 * SCR_TAB);
    if (tmax > sd.tab_tmax) {
        /* the tables depend on t only: a longer prefix replaces a shorter one */
        const int32_t n = tmax + 1;
        std::vector<double> hi(n), lo(n), f(n);
        gepi_build_tables(n, hi.data(), lo.data(), f.data());
        HIP_TRY(hipMemcpyAsync(lhi, hi.data(), sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipMemcpyAsync(llo, lo.data(), sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipMemcpyAsync(lf, f.data(), sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
        sd.tab_tmax = tmax;
    }
    b.lhi = lhi;
    b.llo = llo;
    b.lfact = lf;
    b.state = reinterpret_cast<ScanState *>(base + SCR_STATE);
    b.small = reinterpret_cast<uint64_t *>(base + SCR_SMALL);
    size_t off = SCR_DYN;
    b.cw = reinterpret_cast<CompressWord *>(base + off);
    off += align256(cw_bytes);
    b.clsk = reinterpret_cast<uint64_t *>(base + off);
    off += align256(clsk_bytes);
    b.pend = reinterpret_cast<Pend *>(base + off);
    off += align256(sizeof(Pend) * pcap);
    b.cand[0] = reinterpret_cast<Cand *>(base + off);
    off += align256(sizeof(Cand) * cap);
    b.cand[1] = reinterpret_cast<Cand *>(base + off);
    return GEPI_OK;
}

/* Top-k maintenance by one workgroup (k_topk_one) when the candidates are
 * few enough for one CU to pass over them several times quickly; the
 * multi-workgroup select chain otherwise, and on the paths whose tile
 * numbering k_topk_one does not know (it drops overflowed bands by fp4
 * tiles).                                                                  */
constexpr unsigned long long WG_SELECT_MAX = 1ull << 18;
inline bool wg_select(int path, unsigned long long count) {
    return (path == 3 || path == 4) && count <= WG_SELECT_MAX;
}

TopkDesc topk_desc(const ScanBufs &b, int cur, unsigned long long cap, int64_t k,
                   const gepi_params *par, int64_t t0, int64_t t1, int64_t n_a4) {
    TopkDesc d;
    memset(&d, 0, sizeof d);
    d.st = b.state;
    d.cand = b.cand[cur];
    d.cap = cap;
    d.drop_t0 = t0;
    d.drop_t1 = t1;
    d.report = nullptr;
    d.k = (uint32_t)k;
    d.ffun = par->ffun;
    d.n_a4 = (int32_t)n_a4;
    return d;
}

/* the calling thread's progress hook (gepi_set_progress) */
thread_local gepi_progress_fn g_progress_fn = nullptr;
thread_local void *g_progress_user = nullptr;

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

int run_scan(gepi_panel *p, const uint64_t *mask_ctrl, const uint64_t *mask_case,
             const gepi_params *par, gepi_result *tops, int64_t *n_tops,
             gepi_scan_stats *stats, bool noalloc) {
    auto t_start = std::chrono::steady_clock::now();
    int rc;
    if ((rc = validate_params(par)) != GEPI_OK) return rc;
    if (!n_tops) return gepi_fail(GEPI_ERR_ARG, "null n_tops");
    *n_tops = 0;
    if (par->topmost > 0 && !tops) return gepi_fail(GEPI_ERR_ARG, "null tops");
    if ((rc = set_device(p->device)) != GEPI_OK) return rc;

    gepi_scan_stats st_local;
    memset(&st_local, 0, sizeof st_local);
    const int64_t n_var = p->n_var;
    const int64_t R = 2 * n_var;
    const int64_t total_evals = 5 * (R * (R + 1) / 2);
    st_local.evals = total_evals;

    /* class layout (sample permutation); a sample in both masks counts in both */
    std::vector<int32_t> cls0, cls1;
    for (int32_t s = 0; s < p->n_samples; s++) {
        uint64_t bit = 1ull << (s & 63);
        if (mask_ctrl && (mask_ctrl[s >> 6] & bit)) cls0.push_back(s);
        if (mask_case && (mask_case[s >> 6] & bit)) cls1.push_back(s);
    }
    const int32_t n0 = (int32_t)cls0.size(), n1 = (int32_t)cls1.size();
    st_local.n_class0 = n0;
    st_local.n_class1 = n1;
    if ((int64_t)n0 + n1 > GEPI_MAX_SAMPLES)
        return gepi_fail(GEPI_ERR_LIMIT, "%d class samples exceed the limit of %d", n0 + n1,
                         GEPI_MAX_SAMPLES);

    /* count path: MFMA needs one class of at most MFMA_LOW_MAX samples in the
     * low part of the accumulator; the larger class goes there when it fits. */
    int cls0_high = -1;
    {
        const int big0 = n0 >= n1;
        const int32_t nbig = big0 ? n0 : n1, nsmall = big0 ? n1 : n0;
        if (nbig <= MFMA_LOW_MAX)
            cls0_high = big0 ? 0 : 1;
        else if (nsmall <= MFMA_LOW_MAX)
            cls0_high = big0 ? 1 : 0;
    }
    int path = par->count_path;
    /* fp4: wave tiles from F4W_MIN_KSTEPS K steps on (below, a class's
     * padding K steps, multiplied on the 3-step layout, cost more than the
     * wave tiles save).  Fisher flags about one variant pair in 8; the wave
     * tiles park those in a global list per workgroup, which the
     * allocation-free scan has no room for, so there the LDS-staged kernel
     * (parking in LDS) runs Fisher. */
    if (path == 0) {
        const int32_t ks = (n0 + 127) / 128 + (n1 + 127) / 128;
        const bool fisher = par->ffun != GEPI_FFUNC_ENTROPY;
        path = ks >= F4W_MIN_KSTEPS && !(fisher && noalloc) ? 3 : 4;
    }
    /* Wave tiles: Fisher scans park flagged variant pairs (its bound flags
     * about one pair in eight), entropy scans from F4W_PARK_MIN_KSTEPS K
     * steps, in regions of a global pool, four per CU (at most two
     * workgroups of the kernel are resident on a CU), an eighth of them per
     * XCD.  Parking costs every workgroup a fixed few microseconds (the
     * region claim by atomics, the store wait, two barriers), which a
     * short K loop does not repay: measured, 64 uneven scans of at most
     * about 40 K steps one at a time took 237 ms parking in every band
     * and 197 ms in none (249 ms parking only in bands that let fewer
     * than 0.02 evaluations per variant pair through); the large scan
     * (65 K steps) 126.0 ms parking and 129.2 ms not.                     */
    const uint32_t park_slots = (uint32_t)(((4 * device_cus(p->device)) + 7) / 8 * 8);
    const bool f4w_park =
        path == 3 && !noalloc &&
        (par->ffun != GEPI_FFUNC_ENTROPY || (n0 + 127) / 128 + (n1 + 127) / 128 >= F4W_PARK_MIN_KSTEPS);
    const size_t park_rec_bytes = (size_t)park_slots * F4W_PAIR_CAP * sizeof(PairRec);
    const bool fp4 = path == 3 || path == 4;
    if (noalloc && !fp4)
        return gepi_fail(GEPI_ERR_ARG, "gepi_scan_noalloc runs the fp4 count path only (path %d "
                                       "asked); use gepi_scan", path);
    if (path == 2 && cls0_high < 0)
        return gepi_fail(GEPI_ERR_LIMIT,
                         "MFMA count path needs a class with at most %d samples (have %d and %d)",
                         MFMA_LOW_MAX, n0, n1);
    st_local.count_path = path;

    const int64_t k = par->topmost < total_evals ? par->topmost : total_evals;
    if (k == 0) {
        st_local.ms_scan = ms_since(t_start);
        if (stats) *stats = st_local;
        return GEPI_OK;
    }
    if (k > (int64_t)(1u << 30)) return gepi_fail(GEPI_ERR_LIMIT, "topmost too large");
    const int32_t w0 = ((n0 + 127) / 128) * 4; /* dwords, 128-bit padded */
    const int32_t w1 = ((n1 + 127) / 128) * 4;
    const int32_t wused = w0 + w1;
    const int32_t wc = std::max(KC, ((wused + KC - 1) / KC) * KC);
    std::vector<int32_t> perm((size_t)wc * 32, -1);
    for (int32_t i = 0; i < n0; i++) perm[i] = cls0[i];
    for (int32_t i = 0; i < n1; i++) perm[(size_t)w0 * 32 + i] = cls1[i];
    const int64_t n_var_pad = ((n_var + TM - 1) / TM) * TM;
    /* workgroup tiles of the lower triangle: square tv x tv, or 64 x 32 (fp4) */
    const int32_t tv = path == 2 ? TM : TV;
    const int64_t nt_side = n_var_pad / tv;
    const int64_t n_a4 = n_var_pad / TM;
    const int64_t n_tiles = fp4 ? n_a4 * (n_a4 + 1) : nt_side * (nt_side + 1) / 2;
    const int64_t pairs_per_tile = fp4 ? (int64_t)TM * TB4 : (int64_t)tv * tv;
    const int32_t n_ks = fp4 ? wused / 4 : wused / 2;

    /* int8: n_ks 64-sample words per plane row; fp4: n_u 64-byte super-step
     * rows, then the rows' fourth-K-step set bits ([2][n_var_pad * 3] int32) */
    const int32_t n_u = path == 3 ? f4_super3(w0 / 4) + f4_super3(w1 / 4)
                                  : f4_super(w0 / 4) + f4_super(w1 / 4);
    const size_t f4_lay = (size_t)std::max(n_u, 1) * n_var_pad * 3 * 64;
    const size_t cls_bytes = !fp4 ? (size_t)n_var_pad * 3 * wc * 4 : 0;
    const size_t clsk_bytes =
        fp4 ? f4_lay + (size_t)2 * n_var_pad * 3 * sizeof(int32_t)
                  : path == 2 ? (size_t)std::max(n_ks, 1) * n_var_pad * 3 * 8 : 0;
    /* fp4 path: per 32-sample word of the panel, the compression of each
     * class (mask bits at or beyond n_samples are dropped) */
    std::vector<CompressWord> cw;
    if (fp4) {
        int32_t c0n, c1n;
        build_compress_words(mask_ctrl, mask_case, p->n_samples, p->w64, cw, c0n, c1n);
    }
    /* Candidate buffers: between bands at most k records are kept, and a
     * one-tile band admits at most TILE_EVALS_MAX, so cap >= k +
     * TILE_EVALS_MAX lets a band shrunk to one tile always complete.        */
    const unsigned long long cap_min = (unsigned long long)k + TILE_EVALS_MAX + 1024;
    unsigned long long cap = std::max<unsigned long long>(
        cap_min, (unsigned long long)std::max<int64_t>(4 * k, 1 << 23));
    if (const char *env = getenv("GEPI_CAND_CAP")) {
        /* capacity override for exercising the overflow recovery in tests */
        cap = std::max<unsigned long long>((unsigned long long)std::max(atoll(env), 0ll), cap_min);
    }
    unsigned long long pcap = cap;
    if (const char *env = getenv("GEPI_PEND_CAP")) {
        /* pending-list capacity override for tests: a list larger than the
         * candidate buffers lets one band admit more candidates than fit,
         * which exercises the candidate overflow recovery                 */
        pcap = std::max<unsigned long long>((unsigned long long)std::max(atoll(env), 0ll),
                                            (unsigned long long)TILE_EVALS_MAX + 1024);
    }
    ScanBufs b;
    if (noalloc) {
        /* the smallest lists that always complete: more, smaller bands */
        cap = pcap = cap_min;
        if ((rc = bind_scratch(p, b, n0 + n1 + 1, cw.size() * sizeof(CompressWord), clsk_bytes,
                               cap, pcap)) != GEPI_OK)
            return rc;
    } else {
        /* device bytes allocated by this call: buffers that grow */
        int64_t &alloc = st_local.bytes_allocated;
        if (n0 + n1 + 1 > p->tabs.tmax) alloc += (int64_t)3 * (n0 + n1 + 2) * sizeof(double);
        if ((rc = ensure_tables(p->tabs, n0 + n1 + 1, p->stream)) != GEPI_OK) return rc;
        auto grown = [&alloc](const void *ptr, size_t have, size_t need) {
            if (need > have || !ptr) alloc += (int64_t)need;
        };
        if (cls_bytes) grown(p->d_cls, p->cls_bytes, cls_bytes);
        if (cls_bytes && (rc = grow(&p->d_cls, &p->cls_bytes, cls_bytes)) != GEPI_OK) return rc;
        if (clsk_bytes) grown(p->d_clsk, p->clsk_bytes, clsk_bytes);
        if (clsk_bytes && (rc = grow(&p->d_clsk, &p->clsk_bytes, clsk_bytes)) != GEPI_OK)
            return rc;
        if (!fp4) grown(p->d_perm, p->perm_bytes, perm.size() * 4);
        if (!fp4 && (rc = grow(&p->d_perm, &p->perm_bytes, perm.size() * 4)) != GEPI_OK)
            return rc;
        const size_t park_bytes = f4w_park ? park_rec_bytes + (park_slots + 8) * sizeof(uint32_t) : 0;
        if (park_bytes) grown(p->d_park, p->park_bytes, park_bytes);
        if (park_bytes && (rc = grow(&p->d_park, &p->park_bytes, park_bytes)) != GEPI_OK)
            return rc;
        if (fp4) grown(p->d_cw, p->cw_bytes, cw.size() * sizeof(CompressWord));
        if (fp4 &&
            (rc = grow(&p->d_cw, &p->cw_bytes, cw.size() * sizeof(CompressWord))) != GEPI_OK)
            return rc;
        if (cap > p->buf_cap) alloc += (int64_t)(2 * sizeof(Cand) * cap);
        if (pcap > p->pend_cap) alloc += (int64_t)(sizeof(Pend) * pcap);
        if (cap > p->buf_cap) {
            hipFree(p->d_cand[0]);
            hipFree(p->d_cand[1]);
            p->d_cand[0] = p->d_cand[1] = nullptr;
            p->buf_cap = 0;
            HIP_TRY(hipMalloc(&p->d_cand[0], sizeof(Cand) * cap));
            HIP_TRY(hipMalloc(&p->d_cand[1], sizeof(Cand) * cap));
            p->buf_cap = cap;
        }
        if (pcap > p->pend_cap) {
            hipFree(p->d_pend);
            p->d_pend = nullptr;
            p->pend_cap = 0;
            HIP_TRY(hipMalloc(&p->d_pend, sizeof(Pend) * pcap));
            p->pend_cap = pcap;
        }
        b.cls = p->d_cls;
        b.clsk = p->d_clsk;
        b.perm = p->d_perm;
        b.cw = p->d_cw;
        b.cand[0] = p->d_cand[0];
        b.cand[1] = p->d_cand[1];
        b.pend = p->d_pend;
        b.state = p->d_state;
        b.small = p->d_small;
        b.lhi = p->tabs.lhi;
        b.llo = p->tabs.llo;
        b.lfact = p->tabs.lfact;
    }
    int32_t *d_pc3 = fp4 ? reinterpret_cast<int32_t *>(
                                     reinterpret_cast<char *>(b.clsk) + f4_lay)
                               : nullptr;

    if (fp4)
        HIP_TRY(hipMemcpyAsync(b.cw, cw.data(), cw.size() * sizeof(CompressWord),
                               hipMemcpyHostToDevice, p->stream));
    else
        HIP_TRY(hipMemcpyAsync(b.perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice,
                               p->stream));
    HIP_TRY(hipEventRecord(p->ev0, p->stream));
    if (fp4) {
        const int64_t rows = n_var_pad * 3;
        const unsigned cgrid = (unsigned)((rows + 255) / 256);
        if (n_ks > 0 && path == 3) {
            k_class_f4<3, true><<<cgrid, 256, 0, p->stream>>>(
                p->d_nat, (int32_t)n_var, p->w64, b.cw, rows, f4_super3(w0 / 4),
                reinterpret_cast<uint4 *>(b.clsk), d_pc3);
            HIP_TRY(hipGetLastError());
        } else if (n_ks > 0) {
            k_class_f4<4><<<cgrid, 256, 0, p->stream>>>(
                p->d_nat, (int32_t)n_var, p->w64, b.cw, rows, f4_super(w0 / 4),
                reinterpret_cast<uint4 *>(b.clsk), d_pc3);
            HIP_TRY(hipGetLastError());
        }
    } else {
        size_t lds = (size_t)3 * (p->w64 > 0 ? p->w64 : 1) * 8;
        k_gather_class<<<(unsigned)n_var_pad, 256, lds, p->stream>>>(
            p->d_nat, (int32_t)n_var, p->w64, b.perm, wc, wused, b.cls);
        HIP_TRY(hipGetLastError());
        if (path == 2 && n_ks > 0) {
            dim3 grid((unsigned)(n_var_pad / 32), (unsigned)((n_ks + 31) / 32));
            k_transpose_k<<<grid, 256, 0, p->stream>>>(b.cls, (int32_t)n_var_pad, wc, n_ks,
                                                       b.clsk);
            HIP_TRY(hipGetLastError());
        }

    }
    HIP_TRY(hipEventRecord(p->ev1, p->stream));

    ScanState init;
    memset(&init, 0, sizeof init);
    init.thr_score = INFINITY;
    init.thr_idx = ~0ull;
    HIP_TRY(hipMemcpyAsync(b.state, &init, sizeof init, hipMemcpyHostToDevice, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    {
        float ms = 0;
        hipEventElapsedTime(&ms, p->ev0, p->ev1);
        st_local.ms_prepare = ms;
    }

    int cur = 0; /* active candidate buffer */
    CountArgs ca;
    memset(&ca, 0, sizeof ca);
    ca.planes = b.cls;
    ca.planes_k = b.clsk;
    ca.pc3 = d_pc3;
    ca.n_var = (int32_t)n_var;
    ca.n_var_pad = (int32_t)n_var_pad;
    ca.wc = wc;
    ca.w0 = w0;
    ca.wused = wused;
    ca.n_ks = n_ks;
    ca.ks0 = fp4 ? w0 / 4 : w0 / 2;
    ca.cls0_high = cls0_high > 0 ? 1 : 0;
    ca.f32 = f32_counts(n0, n1) ? 1 : 0;
    ca.n0 = n0;
    ca.n1 = n1;
    ca.st = b.state;
    ca.pend = b.pend;
    ca.pcap = pcap;
    ca.park = f4w_park ? p->d_park : nullptr;
    ca.park_busy = f4w_park ? reinterpret_cast<uint32_t *>(
                                  reinterpret_cast<char *>(p->d_park) + park_rec_bytes)
                            : nullptr;
    ca.park_cap = F4W_PAIR_CAP;
    ca.park_slots = park_slots;
    if (f4w_park)
        HIP_TRY(hipMemsetAsync(ca.park_busy, 0, (park_slots + 8) * sizeof(uint32_t), p->stream));
    ca.lfact = b.lfact;
    ExactArgs ea;
    ea.st = b.state;
    ea.pend = b.pend;
    ea.pcap = pcap;
    ea.cap = cap;
    ea.lhi = b.lhi;
    ea.llo = b.llo;

    const int64_t evals_per_tile = pairs_per_tile * 20;
    /* Bands grow 4x per band up to 8e9 evaluations (195,000 fp4 tiles): each
     * band costs a threshold selection and a host round trip (~0.2 ms), and
     * at the final threshold a band's survivors stay far below the pending
     * list (the next band is also sized from the measured survivor rate). */
    int64_t max_band = std::max<int64_t>(64, (int64_t)(8.0e9 / (double)evals_per_tile));
    if (par->max_grid_area > 0)
        max_band = std::max<int64_t>(1, ((int64_t)par->max_grid_area << 20) / evals_per_tile);
    int64_t band = std::max<int64_t>(1, std::min<int64_t>(max_band, (2 * k) / evals_per_tile + 1));
    /* band-size override for exercising the overflow recovery in tests: the
     * first band and the growth limit, ignoring the survivor-rate estimate */
    int64_t forced_band = 0;
    if (const char *env = getenv("GEPI_BAND_TILES")) {
        forced_band = std::max<long long>(1, atoll(env));
        band = forced_band;
    }
    double ms_count = 0.0, ms_top = 0.0;
    int64_t t0 = 0;
    unsigned long long count = 0, exact_total = 0;
    int bands = 0;
    const bool entropy = par->ffun == GEPI_FFUNC_ENTROPY;
    /* fp4: the filter runs on the fp32 accumulators when every product of a
     * class-0 and a class-1 count is exact in float */
    const bool f32 = ca.f32 != 0;
    /* wave tiles: the kernel built with parking where the scan parks */
    const bool park = f4w_park;
    while (t0 < n_tiles) {
        const int64_t t1 = std::min<int64_t>(n_tiles, t0 + band);
        ca.tile_begin = t0;
        ca.n_tiles = (int32_t)(t1 - t0);
        ea.cand = b.cand[cur];
        k_reset_pending<<<1, 1, 0, p->stream>>>(b.state);
        HIP_TRY(hipEventRecord(p->ev0, p->stream));
        const unsigned grid = path == 3 ? (unsigned)((t1 - t0 + 1) / 2) : (unsigned)(t1 - t0);
        if (path == 3) {
            if (entropy && f32 && park)
                k_count_f4w<GEPI_FFUNC_ENTROPY, true, true><<<grid, F4W_THREADS, 0, p->stream>>>(ca);
            else if (entropy && f32)
                k_count_f4w<GEPI_FFUNC_ENTROPY, true, false><<<grid, F4W_THREADS, 0, p->stream>>>(ca);
            else if (entropy && park)
                k_count_f4w<GEPI_FFUNC_ENTROPY, false, true><<<grid, F4W_THREADS, 0, p->stream>>>(ca);
            else if (entropy)
                k_count_f4w<GEPI_FFUNC_ENTROPY, false, false><<<grid, F4W_THREADS, 0, p->stream>>>(ca);
            else if (f32)
                k_count_f4w<GEPI_FFUNC_FISHERPV, true, true><<<grid, F4W_THREADS, 0, p->stream>>>(ca);
            else
                k_count_f4w<GEPI_FFUNC_FISHERPV, false, true><<<grid, F4W_THREADS, 0, p->stream>>>(ca);
        } else if (path == 4) {
            if (entropy && f32)
                k_count_f4<GEPI_FFUNC_ENTROPY, true><<<grid, F4_THREADS, 0, p->stream>>>(ca);
            else if (entropy)
                k_count_f4<GEPI_FFUNC_ENTROPY, false><<<grid, F4_THREADS, 0, p->stream>>>(ca);
            else if (f32)
                k_count_f4<GEPI_FFUNC_FISHERPV, true><<<grid, F4_THREADS, 0, p->stream>>>(ca);
            else
                k_count_f4<GEPI_FFUNC_FISHERPV, false><<<grid, F4_THREADS, 0, p->stream>>>(ca);
        } else if (path == 2) {
            if (entropy)
                k_count_mfma<GEPI_FFUNC_ENTROPY><<<grid, MF_THREADS, 0, p->stream>>>(ca);
            else
                k_count_mfma<GEPI_FFUNC_FISHERPV><<<grid, MF_THREADS, 0, p->stream>>>(ca);
        } else {
            if (entropy)
                k_count_popc<GEPI_FFUNC_ENTROPY><<<grid, POPC_THREADS, 0, p->stream>>>(ca);
            else
                k_count_popc<GEPI_FFUNC_FISHERPV><<<grid, POPC_THREADS, 0, p->stream>>>(ca);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(p->ev1, p->stream));
        auto ts = std::chrono::steady_clock::now();
        if (entropy)
            k_exact<GEPI_FFUNC_ENTROPY><<<1024, 256, 0, p->stream>>>(ea);
        else
            k_exact<GEPI_FFUNC_FISHERPV><<<1024, 256, 0, p->stream>>>(ea);
        HIP_TRY(hipGetLastError());
        ScanState hs;
        if ((rc = read_state(p, b, &hs)) != GEPI_OK) return rc;
        float ms = 0;
        hipEventElapsedTime(&ms, p->ev0, p->ev1);
        ms_count += ms;
        bands++;
        if (hs.pend_overflow) {
            /* the pending list lost survivors; the candidates are untouched
             * (k_exact skipped), so redo the band in smaller pieces        */
            st_local.pend_overflows++;
            band = std::max<int64_t>(1, (t1 - t0) / 4);
            ms_top += ms_since(ts);
            continue;
        }
        exact_total += hs.pend_count;
        const unsigned long long added = hs.count - count;
        if (hs.overflow) {
            /* candidates lost: threshold from what was kept, drop this band's
             * records and redo the band                                     */
            st_local.cand_overflows++;
            const unsigned long long kept = std::min<unsigned long long>(hs.count, cap);
            if (wg_select(path, kept)) {
                /* reads the overflow flag, clamps the count to cap and drops
                 * the band's records itself */
                k_topk_one<<<1, TOPK_THREADS, 0, p->stream>>>(topk_desc(b, cur, cap, k, par, t0, t1, n_a4));
            } else {
                hs.count = kept;
                hs.overflow = 0;
                HIP_TRY(hipMemcpyAsync(b.state, &hs, sizeof hs, hipMemcpyHostToDevice, p->stream));
                launch_select(p, b, b.cand[cur], (uint32_t)k, kept);
                k_compact<<<4096, 256, 0, p->stream>>>(b.cand[cur], b.state, b.cand[cur ^ 1],
                                                       t0, t1, tv, fp4 ? (int32_t)n_a4 : 0);
                k_finalize<<<1, 1, 0, p->stream>>>(b.state, par->ffun);
                cur ^= 1;
            }
            HIP_TRY(hipGetLastError());
            ScanState h2;
            if ((rc = read_state(p, b, &h2)) != GEPI_OK) return rc;
            count = h2.count;
            exact_total -= hs.pend_count;
            band = std::max<int64_t>(1, (t1 - t0) / 4);
            ms_top += ms_since(ts);
            continue;
        }
        if (hs.count > (unsigned long long)k) {
            if (wg_select(path, hs.count)) {
                k_topk_one<<<1, TOPK_THREADS, 0, p->stream>>>(topk_desc(b, cur, cap, k, par, 0, 0, n_a4));
            } else {
                launch_select(p, b, b.cand[cur], (uint32_t)k, hs.count);
                k_compact<<<4096, 256, 0, p->stream>>>(b.cand[cur], b.state, b.cand[cur ^ 1],
                                                       0, 0, tv, fp4 ? (int32_t)n_a4 : 0);
                k_finalize<<<1, 1, 0, p->stream>>>(b.state, par->ffun);
                cur ^= 1;
            }
            HIP_TRY(hipGetLastError());
            count = (unsigned long long)k;
        } else {
            count = hs.count;
        }
        ms_top += ms_since(ts);
        /* size the next band so that its filter survivors fit comfortably */
        const double rate = (double)std::max<unsigned long long>(hs.pend_count, added) /
                            (double)(t1 - t0);
        const int64_t room = (int64_t)(cap - (unsigned long long)k) / 2;
        const int64_t next = rate > 0.0 ? (int64_t)((double)room / rate) : max_band;
        band = std::max<int64_t>(1, std::min<int64_t>({max_band, next, band * 4}));
        if (forced_band) band = std::min<int64_t>(forced_band, (t1 - t0) * 4);
        t0 = t1;
        if (par->intermediate && g_progress_fn) {
            /* the records held so far, sorted, for the caller's report */
            std::vector<Cand> held((size_t)count);
            if (count > 0)
                HIP_TRY(hipMemcpyAsync(held.data(), b.cand[cur], sizeof(Cand) * count,
                                       hipMemcpyDeviceToHost, p->stream));
            HIP_TRY(hipStreamSynchronize(p->stream));
            std::vector<gepi_result> cur_tops((size_t)std::min<int64_t>(k, (int64_t)count));
            gepi_progress pg;
            memset(&pg, 0, sizeof pg);
            pg.bands = bands;
            pg.tiles_done = t0;
            pg.tiles_total = n_tiles;
            pg.ms_elapsed = ms_since(t_start);
            pg.ms_remaining = pg.ms_elapsed * (double)(n_tiles - t0) / (double)t0;
            pg.n_tops = candidates_to_results(held, k, cur_tops.data());
            pg.tops = cur_tops.data();
            g_progress_fn(&pg, g_progress_user);
        }
    }
    /* final extraction: exactly k records, sorted by key on the host */
    auto ts = std::chrono::steady_clock::now();
    std::vector<Cand> host((size_t)count);
    if (count > 0)
        HIP_TRY(hipMemcpyAsync(host.data(), b.cand[cur], sizeof(Cand) * count,
                               hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    const int64_t n = candidates_to_results(host, k, tops);
    *n_tops = n;
    ms_top += ms_since(ts);
    st_local.ms_count_kernels = ms_count;
    st_local.ms_select = ms_top;
    st_local.exact_evals = (int64_t)exact_total;
    st_local.bands = bands;
    st_local.ms_scan = ms_since(t_start);
    if (n != k)
        return gepi_fail(GEPI_ERR_INTERNAL, "top-k extraction returned %lld of %lld records",
                         (long long)n, (long long)k);
    if (stats) *stats = st_local;
    return GEPI_OK;
}

int upload_planes(gepi_panel *p, const uint64_t *bits, const uint64_t *valid, int64_t n_rows,
                  int32_t words) {
    const int64_t n_var = n_rows / 2;
    std::vector<uint64_t> nat((size_t)(n_var > 0 ? n_var : 1) * 3 * (words > 0 ? words : 1), 0);
    for (int64_t v = 0; v < n_var; v++) {
        const uint64_t *b1 = bits + (size_t)(2 * v) * words, *b2 = b1 + words;
        const uint64_t *v1 = valid + (size_t)(2 * v) * words;
        uint64_t *o = nat.data() + (size_t)v * 3 * words;
        for (int32_t w = 0; w < words; w++) {
            o[w] = b1[w] & v1[w];
            o[words + w] = b2[w] & v1[w];
            o[2 * words + w] = v1[w];
        }
    }
    /* on the panel's stream, after panel_alloc's zero fill of the same
     * buffer (a null-stream copy would not be ordered with that non-blocking
     * stream); nat lives until the synchronisation */
    if (n_var > 0 && words > 0) {
        HIP_TRY(hipMemcpyAsync(p->d_nat, nat.data(), nat.size() * 8, hipMemcpyHostToDevice,
                               p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
    }
    return GEPI_OK;
}

} // namespace

/* ------------------------------------------------------------------------ */
/* C ABI                                                                     */

extern "C" void gepi_params_init(gepi_params *par) {
    if (!par) return;
    memset(par, 0, sizeof *par);
    par->gpu_num = 0;
    par->ffun = GEPI_FFUNC_ENTROPY;
    par->topmost = GEPI_DEFAULT_TOPMOST;
    par->count_path = 0;
    par->blk_x = par->blk_y = 16;
    par->max_grid_area = 0;
}

extern "C" int gepi_set_progress(gepi_progress_fn fn, void *user) {
    g_progress_fn = fn;
    g_progress_user = fn ? user : nullptr;
    return GEPI_OK;
}

extern "C" int gepi_device_count(int32_t *count) {
    if (!count) return gepi_fail(GEPI_ERR_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return GEPI_OK;
}

extern "C" int gepi_panel_create(const uint64_t *bit_data, const uint64_t *valid_data,
                                 int64_t n_rows, int32_t n_samples, int32_t device,
                                 gepi_panel **out) {
    if (!out) return gepi_fail(GEPI_ERR_ARG, "null out");
    *out = nullptr;
    if (n_rows < 0 || n_samples < 0) return gepi_fail(GEPI_ERR_ARG, "negative size");
    if (n_rows % 2) return gepi_fail(GEPI_ERR_LAYOUT, "odd row count %lld", (long long)n_rows);
    if (n_rows >= (1LL << 29)) return gepi_fail(GEPI_ERR_LIMIT, "too many rows");
    const int32_t words = (n_samples + 63) / 64;
    if (n_rows > 0 && words > 0 && (!bit_data || !valid_data))
        return gepi_fail(GEPI_ERR_ARG, "null plane data");
    int rc;
    if (n_rows > 0 && words > 0 && (rc = check_pairing(valid_data, n_rows, words)) != GEPI_OK)
        return rc;
    if ((rc = set_device(device)) != GEPI_OK) return rc;
    gepi_panel *p = new gepi_panel();
    if ((rc = panel_alloc(p, device, n_rows / 2, n_samples)) != GEPI_OK ||
        (rc = upload_planes(p, bit_data, valid_data, n_rows, words)) != GEPI_OK) {
        panel_free(p);
        return rc;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) {
        panel_free(p);
        return gepi_fail(GEPI_ERR_DEVICE, "panel upload failed");
    }
    *out = p;
    return GEPI_OK;
}

extern "C" int gepi_panel_create_bed(const uint8_t *bed_payload, int64_t n_variants,
                                     int32_t n_samples, int32_t device, gepi_panel **out) {
    if (!out) return gepi_fail(GEPI_ERR_ARG, "null out");
    *out = nullptr;
    if (n_variants < 0 || n_samples < 0) return gepi_fail(GEPI_ERR_ARG, "negative size");
    if (n_variants >= (1LL << 28)) return gepi_fail(GEPI_ERR_LIMIT, "too many variants");
    const int64_t bps = ((int64_t)n_samples + 3) / 4;
    if (n_variants > 0 && bps > 0 && !bed_payload) return gepi_fail(GEPI_ERR_ARG, "null payload");
    int rc;
    if ((rc = set_device(device)) != GEPI_OK) return rc;
    gepi_panel *p = new gepi_panel();
    if ((rc = panel_alloc(p, device, n_variants, n_samples)) != GEPI_OK) {
        panel_free(p);
        return rc;
    }
    const int64_t bytes = n_variants * bps;
    if (bytes > 0 && p->w64 > 0) {
        uint8_t *d_bed = nullptr;
        if (hipMalloc(&d_bed, (size_t)bytes) != hipSuccess) {
            panel_free(p);
            return gepi_fail(GEPI_ERR_NOMEM, "bed staging allocation failed");
        }
        hipMemcpyAsync(d_bed, bed_payload, (size_t)bytes, hipMemcpyHostToDevice, p->stream);
        int64_t threads = n_variants * p->w64;
        k_decode_bed<<<(unsigned)((threads + 255) / 256), 256, 0, p->stream>>>(
            d_bed, n_variants, n_samples, p->w64, bps, p->d_nat);
        hipError_t e = hipStreamSynchronize(p->stream);
        hipError_t e2 = hipGetLastError();
        hipFree(d_bed);
        if (e != hipSuccess || e2 != hipSuccess) {
            panel_free(p);
            return gepi_fail(GEPI_ERR_DEVICE, "bed decode failed: %s",
                             hipGetErrorString(e != hipSuccess ? e : e2));
        }
    }
    *out = p;
    return GEPI_OK;
}

extern "C" int gepi_panel_destroy(gepi_panel *panel) {
    if (!panel) return gepi_fail(GEPI_ERR_ARG, "null panel");
    panel_free(panel);
    return GEPI_OK;
}

namespace {

/* Missing-call counts of a resident panel: per variant (n_var) and per
 * sample (n_samples), on the panel's device and stream. */
int panel_missing(gepi_panel *p, std::vector<int32_t> &var_miss, std::vector<int32_t> &samp_miss) {
    var_miss.assign((size_t)p->n_var, 0);
    samp_miss.assign((size_t)p->n_samples, 0);
    if (p->n_var == 0 || p->n_samples == 0) {
        for (auto &m : var_miss) m = p->n_samples;
        for (auto &m : samp_miss) m = (int32_t)p->n_var;
        return GEPI_OK;
    }
    int32_t *d_vm = nullptr, *d_sv = nullptr;
    HIP_TRY(hipMalloc(&d_vm, sizeof(int32_t) * (size_t)p->n_var));
    if (hipMalloc(&d_sv, sizeof(int32_t) * (size_t)p->n_samples) != hipSuccess) {
        hipFree(d_vm);
        return gepi_fail(GEPI_ERR_NOMEM, "QA buffers");
    }
    hipMemsetAsync(d_sv, 0, sizeof(int32_t) * (size_t)p->n_samples, p->stream);
    k_qa_variants<<<(unsigned)((p->n_var + 3) / 4), 256, 0, p->stream>>>(p->d_nat, p->n_var, p->w64,
                                                                        p->n_samples, d_vm);
    const int32_t chunk = 1024;
    dim3 grid((unsigned)((p->n_samples + 255) / 256), (unsigned)((p->n_var + chunk - 1) / chunk));
    k_qa_samples<<<grid, 256, 0, p->stream>>>(p->d_nat, p->n_var, p->w64, p->n_samples, chunk, d_sv);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(var_miss.data(), d_vm, sizeof(int32_t) * (size_t)p->n_var,
                           hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(samp_miss.data(), d_sv, sizeof(int32_t) * (size_t)p->n_samples,
                           hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    hipFree(d_vm);
    hipFree(d_sv);
    if (e != hipSuccess)
        return gepi_fail(GEPI_ERR_DEVICE, "QA kernels failed: %s", hipGetErrorString(e));
    for (auto &m : samp_miss) m = (int32_t)p->n_var - m; /* valid calls -> missing */
    return GEPI_OK;
}

/* The reference's tests (src/reader/PlinkReader.h:318-319 per variant,
 * :281-282 per sample), which count allele rows: both rows of a variant share
 * its calls, so a sample misses 2 rows per missing variant out of 2*n_var. */
inline bool variant_bad(int32_t miss, int32_t n_samples, float thr) {
    return n_samples > 0 && (float)miss / (float)n_samples > thr;
}
inline bool sample_bad(int32_t miss, int64_t n_var, float thr) {
    return n_var > 0 && (float)(2 * (int64_t)miss) / (float)(2 * n_var) > thr;
}

} // namespace

extern "C" int gepi_panel_qa(gepi_panel *p, float max_miss_snps, float max_miss_samples,
                             int32_t *variant_missing, int32_t *sample_missing,
                             int64_t *n_bad_variants, int64_t *n_bad_samples) {
    if (!p) return gepi_fail(GEPI_ERR_ARG, "null panel");
    int rc;
    if ((rc = set_device(p->device)) != GEPI_OK) return rc;
    std::vector<int32_t> vm, sm;
    if ((rc = panel_missing(p, vm, sm)) != GEPI_OK) return rc;
    int64_t bv = 0, bs = 0;
    for (int32_t m : vm) bv += variant_bad(m, p->n_samples, max_miss_snps) ? 1 : 0;
    for (int32_t m : sm) bs += sample_bad(m, p->n_var, max_miss_samples) ? 1 : 0;
    if (variant_missing && !vm.empty()) memcpy(variant_missing, vm.data(), sizeof(int32_t) * vm.size());
    if (sample_missing && !sm.empty()) memcpy(sample_missing, sm.data(), sizeof(int32_t) * sm.size());
    if (n_bad_variants) *n_bad_variants = bv;
    if (n_bad_samples) *n_bad_samples = bs;
    return GEPI_OK;
}

extern "C" int gepi_panel_filter(gepi_panel *p, float max_miss_snps, float max_miss_samples,
                                 gepi_panel **out, int64_t *kept_variants, int64_t *n_kept,
                                 uint64_t *keep_samples) {
    if (!p || !out) return gepi_fail(GEPI_ERR_ARG, "null argument");
    *out = nullptr;
    int rc;
    if ((rc = set_device(p->device)) != GEPI_OK) return rc;
    std::vector<int32_t> vm, sm;
    if ((rc = panel_missing(p, vm, sm)) != GEPI_OK) return rc;
    std::vector<int64_t> keep;
    for (int64_t v = 0; v < p->n_var; v++)
        if (!variant_bad(vm[(size_t)v], p->n_samples, max_miss_snps)) keep.push_back(v);
    if (keep_samples) {
        memset(keep_samples, 0, sizeof(uint64_t) * (size_t)p->w64);
        for (int32_t s = 0; s < p->n_samples; s++)
            if (!sample_bad(sm[(size_t)s], p->n_var, max_miss_samples))
                keep_samples[s >> 6] |= 1ull << (s & 63);
    }
    gepi_panel *q = new gepi_panel();
    if ((rc = panel_alloc(q, p->device, (int64_t)keep.size(), p->n_samples)) != GEPI_OK) {
        panel_free(q);
        return rc;
    }
    if (!keep.empty() && p->w64 > 0) {
        int64_t *d_idx = nullptr;
        if (hipMalloc(&d_idx, sizeof(int64_t) * keep.size()) != hipSuccess) {
            panel_free(q);
            return gepi_fail(GEPI_ERR_NOMEM, "filter index");
        }
        hipMemcpyAsync(d_idx, keep.data(), sizeof(int64_t) * keep.size(), hipMemcpyHostToDevice,
                       q->stream);
        const int64_t threads = (int64_t)keep.size() * 3 * p->w64;
        /* the source panel's pending work (its own stream) is complete: the
         * QA above synchronised it */
        k_gather_variants<<<(unsigned)((threads + 255) / 256), 256, 0, q->stream>>>(
            p->d_nat, d_idx, (int64_t)keep.size(), p->w64, q->d_nat);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(q->stream);
        hipFree(d_idx);
        if (e != hipSuccess) {
            panel_free(q);
            return gepi_fail(GEPI_ERR_DEVICE, "variant gather failed: %s", hipGetErrorString(e));
        }
    }
    if (kept_variants && !keep.empty()) memcpy(kept_variants, keep.data(), sizeof(int64_t) * keep.size());
    if (n_kept) *n_kept = (int64_t)keep.size();
    *out = q;
    return GEPI_OK;
}

extern "C" int gepi_panel_download(gepi_panel *p, uint64_t *bit_data, uint64_t *valid_data) {
    if (!p || !bit_data || !valid_data) return gepi_fail(GEPI_ERR_ARG, "null argument");
    int rc;
    if ((rc = set_device(p->device)) != GEPI_OK) return rc;
    const int32_t W = p->w64;
    std::vector<uint64_t> nat((size_t)p->n_var * 3 * W);
    if (!nat.empty()) {
        /* ordered after the panel's own queued work (its non-blocking stream) */
        HIP_TRY(hipMemcpyAsync(nat.data(), p->d_nat, nat.size() * 8, hipMemcpyDeviceToHost,
                               p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
    }
    for (int64_t v = 0; v < p->n_var; v++)
        for (int32_t w = 0; w < W; w++) {
            const uint64_t *o = nat.data() + (size_t)v * 3 * W;
            bit_data[(size_t)(2 * v) * W + w] = o[w];
            bit_data[(size_t)(2 * v + 1) * W + w] = o[W + w];
            valid_data[(size_t)(2 * v) * W + w] = o[2 * W + w];
            valid_data[(size_t)(2 * v + 1) * W + w] = o[2 * W + w];
        }
    return GEPI_OK;
}

extern "C" int gepi_scan(gepi_panel *panel, const uint64_t *mask_ctrl, const uint64_t *mask_case,
                         const gepi_params *par, gepi_result *tops, int64_t *n_tops,
                         gepi_scan_stats *stats) {
    if (!panel) return gepi_fail(GEPI_ERR_ARG, "null panel");
    return run_scan(panel, mask_ctrl, mask_case, par, tops, n_tops, stats, false);
}

extern "C" int gepi_scan_noalloc(gepi_panel *panel, const uint64_t *mask_ctrl,
                                 const uint64_t *mask_case, const gepi_params *par,
                                 gepi_result *tops, int64_t *n_tops, gepi_scan_stats *stats) {
    if (!panel) return gepi_fail(GEPI_ERR_ARG, "null panel");
    return run_scan(panel, mask_ctrl, mask_case, par, tops, n_tops, stats, true);
}

extern "C" int gepi_process_gpu_1(const uint64_t *bit_data, const uint64_t *valid_data,
                                  const uint64_t *mask_ctrl, const uint64_t *mask_case,
                                  int64_t n_rows, int32_t n_samples, const gepi_params *par,
                                  gepi_result *tops, int64_t *n_tops, gepi_scan_stats *stats) {
    int rc;
    if ((rc = validate_params(par)) != GEPI_OK) return rc;
    gepi_panel *p = nullptr;
    if ((rc = gepi_panel_create(bit_data, valid_data, n_rows, n_samples, par->gpu_num, &p)) !=
        GEPI_OK)
        return rc;
    rc = run_scan(p, mask_ctrl, mask_case, par, tops, n_tops, stats, false);
    panel_free(p);
    return rc;
}

/* One table scored on the host by the exact stage's own code (gepi_score.h,
 * built here by the same compiler and flags as the device code). */
extern "C" int gepi_score_table(int32_t w, int32_t x, int32_t y, int32_t z, int32_t ffun,
                                double *score) {
    if (!score) return gepi_fail(GEPI_ERR_ARG, "score is NULL");
    if (w < 0 || x < 0 || w > y || x > z)
        return gepi_fail(GEPI_ERR_ARG, "(w, x, y, z) = (%d, %d, %d, %d) is not a table", w, x, y, z);
    if ((int64_t)y + z > GEPI_MAX_SAMPLES)
        return gepi_fail(GEPI_ERR_LIMIT, "%lld samples exceed the limit of %d", (long long)y + z,
                         GEPI_MAX_SAMPLES);
    if (ffun == GEPI_FFUNC_FISHERPV) {
        *score = gepi::fisher_exact(w, x, y, z);
        return GEPI_OK;
    }
    if (ffun != GEPI_FFUNC_ENTROPY) return gepi_fail(GEPI_ERR_ARG, "unknown objective %d", ffun);
    /* t log2 t tables, grown to the largest table scored so far */
    static std::mutex mu;
    static std::vector<double> lhi, llo, lf;
    std::lock_guard<std::mutex> lock(mu);
    const size_t n = (size_t)y + z + 1;
    if (lhi.size() < n) {
        const size_t m = std::max<size_t>(n, 2 * lhi.size());
        lhi.assign(m, 0.0);
        llo.assign(m, 0.0);
        lf.assign(m, 0.0);
        gepi_build_tables((int32_t)m, lhi.data(), llo.data(), lf.data());
    }
    *score = gepi::entropy_exact(w, x, y, z, lhi.data(), llo.data());
    return GEPI_OK;
}
