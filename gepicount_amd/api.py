"""ctypes binding of libgepicount.so (include/gepicount.h).

Python mirror of the reference's scan interface:
    process_gpu_1(reader_planes, tops, par)   src/proc_gpu.h:134, src/kernels/kernel1.cu:508-511
plus resident panels (Panel) and the PLINK reader (read_plink).  Everything
runs in the native library; there is no Python or CPU fallback, and loading
fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgepicount.so")

GEPI_FFUNC_ENTROPY = 0
GEPI_FFUNC_FISHERPV = 1
FUNC_NAMES = ("AND", "NAND", "ANDN", "NANDN", "XOR")
STATUS = {0: "GEPI_OK", -1: "GEPI_ERR_ARG", -2: "GEPI_ERR_LAYOUT", -3: "GEPI_ERR_LIMIT",
          -4: "GEPI_ERR_NOMEM", -5: "GEPI_ERR_DEVICE", -6: "GEPI_ERR_IO", -7: "GEPI_ERR_FORMAT",
          -8: "GEPI_ERR_NOCLASS", -9: "GEPI_ERR_INTERNAL"}

RESULT_DTYPE = np.dtype([("sA", "<i8"), ("sB", "<i8"), ("w", "<i4"), ("x", "<i4"),
                         ("y", "<i4"), ("z", "<i4"), ("obj", "<f8"), ("fun", "<i4"),
                         ("reserved", "<i4")])

# Names of every entry point declared in include/gepicount.h.
EXPORTS = ("gepi_params_init", "gepi_process_gpu_1", "gepi_panel_create", "gepi_panel_create_bed",
           "gepi_panel_destroy", "gepi_scan", "gepi_scan_noalloc", "gepi_panel_download",
           "gepi_panel_qa",
           "gepi_panel_filter", "gepi_plink_open",
           "gepi_plink_close", "gepi_plink_check_quality", "gepi_print_interactions",
           "gepi_fisher_pval", "gepi_score_table", "gepi_last_error", "gepi_version",
           "gepi_device_count", "gepi_set_progress",
           "gepi_engine_create", "gepi_engine_destroy", "gepi_engine_scan_batch",
           "gepi_engine_autotune", "gepi_engine_set_tuning",
           "gepi_server_create", "gepi_server_destroy", "gepi_server_add_panel_bed",
           "gepi_server_release_panel", "gepi_server_submit", "gepi_server_done",
           "gepi_server_wait", "gepi_server_get_stats")


class GepiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("gpu_num", ctypes.c_int32), ("ffun", ctypes.c_int32),
                ("topmost", ctypes.c_int64), ("count_path", ctypes.c_int32),
                ("intermediate", ctypes.c_int32), ("blk_x", ctypes.c_int32),
                ("blk_y", ctypes.c_int32), ("max_grid_area", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class ScanStats(ctypes.Structure):
    _fields_ = [("ms_scan", ctypes.c_double), ("ms_count_kernels", ctypes.c_double),
                ("ms_prepare", ctypes.c_double), ("ms_select", ctypes.c_double),
                ("evals", ctypes.c_int64), ("exact_evals", ctypes.c_int64),
                ("bands", ctypes.c_int32), ("count_path", ctypes.c_int32),
                ("n_class0", ctypes.c_int32), ("n_class1", ctypes.c_int32),
                ("pend_overflows", ctypes.c_int32), ("cand_overflows", ctypes.c_int32),
                ("bytes_allocated", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchItem(ctypes.Structure):
    _fields_ = [("panel", ctypes.c_void_p), ("mask_ctrl", ctypes.POINTER(ctypes.c_uint64)),
                ("mask_case", ctypes.POINTER(ctypes.c_uint64)), ("topmost", ctypes.c_int64),
                ("ffun", ctypes.c_int32), ("status", ctypes.c_int32),
                ("tops", ctypes.c_void_p), ("n_tops", ctypes.c_int64), ("stats", ScanStats)]


class BatchStats(ctypes.Structure):
    _fields_ = [("ms_batch", ctypes.c_double), ("ms_prepare", ctypes.c_double),
                ("ms_count_kernels", ctypes.c_double), ("evals", ctypes.c_int64),
                ("rounds", ctypes.c_int32), ("items_failed", ctypes.c_int32),
                ("graph_rounds", ctypes.c_int32), ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneClass(ctypes.Structure):
    _fields_ = [("max_ksteps", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
                ("tile", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("band_evals", ctypes.c_double), ("ms_best", ctypes.c_double),
                ("ms_default", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerParams(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("autotune", ctypes.c_int32),
                ("hbm_budget", ctypes.c_int64), ("round_wg_budget", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("failed", ctypes.c_int64), ("evals", ctypes.c_int64), ("rounds", ctypes.c_int64),
                ("max_active", ctypes.c_int32), ("panels", ctypes.c_int32),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("busy_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PlinkStruct(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int64), ("n_samples", ctypes.c_int32),
                ("words", ctypes.c_int32),
                ("bit_data", ctypes.POINTER(ctypes.c_uint64)),
                ("valid_data", ctypes.POINTER(ctypes.c_uint64)),
                ("mask_ctrl", ctypes.POINTER(ctypes.c_uint64)),
                ("mask_case", ctypes.POINTER(ctypes.c_uint64)),
                ("pheno", ctypes.POINTER(ctypes.c_int32)),
                ("sample_name", ctypes.POINTER(ctypes.c_char_p)),
                ("snp_id", ctypes.POINTER(ctypes.c_char_p)),
                ("allele", ctypes.POINTER(ctypes.c_char)),
                ("chr", ctypes.POINTER(ctypes.c_int32)),
                ("pos", ctypes.POINTER(ctypes.c_int32)),
                ("total_missing", ctypes.c_int64),
                ("bed_payload", ctypes.POINTER(ctypes.c_uint8))]


_lib = None


def lib():
    """Loads libgepicount.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libgepicount.so is not built (run __graft_entry__.build() or "
                           "make -C gepicount_amd/csrc)")
    l = ctypes.CDLL(LIB_PATH)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    vp = ctypes.c_void_p
    l.gepi_params_init.argtypes = [ctypes.POINTER(Params)]
    l.gepi_params_init.restype = None
    l.gepi_process_gpu_1.argtypes = [u64p, u64p, u64p, u64p, ctypes.c_int64, ctypes.c_int32,
                                     ctypes.POINTER(Params), vp, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ScanStats)]
    l.gepi_panel_create.argtypes = [u64p, u64p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.POINTER(vp)]
    l.gepi_panel_create_bed.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(vp)]
    l.gepi_panel_destroy.argtypes = [vp]
    l.gepi_scan.argtypes = [vp, u64p, u64p, ctypes.POINTER(Params), vp,
                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ScanStats)]
    l.gepi_scan_noalloc.argtypes = [vp, u64p, u64p, ctypes.POINTER(Params), vp,
                                    ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ScanStats)]
    l.gepi_panel_download.argtypes = [vp, u64p, u64p]
    i32p, i64p = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)
    l.gepi_panel_qa.argtypes = [vp, ctypes.c_float, ctypes.c_float, i32p, i32p, i64p, i64p]
    l.gepi_panel_filter.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.POINTER(vp), i64p,
                                    i64p, u64p]
    l.gepi_plink_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(PlinkStruct))]
    l.gepi_plink_close.argtypes = [ctypes.POINTER(PlinkStruct)]
    l.gepi_plink_close.restype = None
    l.gepi_plink_check_quality.argtypes = [ctypes.POINTER(PlinkStruct), ctypes.c_float,
                                           ctypes.c_float, ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_int32)]
    l.gepi_print_interactions.argtypes = [ctypes.POINTER(PlinkStruct), vp, ctypes.c_int64,
                                          ctypes.c_int32, ctypes.c_int32]
    l.gepi_fisher_pval.argtypes = [ctypes.c_int32] * 4
    l.gepi_fisher_pval.restype = ctypes.c_float
    l.gepi_last_error.restype = ctypes.c_char_p
    l.gepi_version.restype = ctypes.c_char_p
    l.gepi_device_count.argtypes = [ctypes.POINTER(ctypes.c_int32)]
    l.gepi_set_progress.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    l.gepi_score_table.argtypes = [ctypes.c_int32] * 5 + [ctypes.POINTER(ctypes.c_double)]
    l.gepi_engine_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp)]
    l.gepi_engine_destroy.argtypes = [vp]
    l.gepi_engine_scan_batch.argtypes = [vp, ctypes.POINTER(BatchItem), ctypes.c_int32,
                                         ctypes.POINTER(BatchStats)]
    l.gepi_engine_autotune.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(TuneClass), ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_int32)]
    l.gepi_engine_set_tuning.argtypes = [vp, ctypes.POINTER(TuneClass), ctypes.c_int32]
    u64 = ctypes.c_uint64
    l.gepi_server_create.argtypes = [ctypes.POINTER(ServerParams), ctypes.POINTER(vp)]
    l.gepi_server_destroy.argtypes = [vp]
    l.gepi_server_add_panel_bed.argtypes = [vp, vp, ctypes.c_int64, ctypes.c_int32,
                                            ctypes.POINTER(u64)]
    l.gepi_server_release_panel.argtypes = [vp, u64]
    l.gepi_server_submit.argtypes = [vp, u64, u64p, u64p, ctypes.c_int64, ctypes.c_int32,
                                     ctypes.POINTER(u64)]
    l.gepi_server_done.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int32)]
    l.gepi_server_wait.argtypes = [vp, u64, vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ScanStats)]
    l.gepi_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    for name in EXPORTS:
        if name not in ("gepi_params_init", "gepi_plink_close", "gepi_fisher_pval",
                        "gepi_last_error", "gepi_version"):
            getattr(l, name).restype = ctypes.c_int
    _lib = l
    return l


def _check(rc):
    if rc != 0:
        raise GepiError(rc, lib().gepi_last_error().decode())


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def make_params(topmost=1000000, ffun=GEPI_FFUNC_ENTROPY, gpu_num=0, count_path=0,
                max_grid_area=0):
    p = Params()
    lib().gepi_params_init(ctypes.byref(p))
    p.topmost = int(topmost)
    p.ffun = int(ffun)
    p.gpu_num = int(gpu_num)
    p.count_path = int(count_path)
    p.max_grid_area = int(max_grid_area)
    return p


def device_count():
    n = ctypes.c_int32(0)
    _check(lib().gepi_device_count(ctypes.byref(n)))
    return n.value


def score_table(w, x, y, z, ffun=GEPI_FFUNC_ENTROPY):
    """Exact objective of one table (what a result's obj holds for these
    counts): -I(F;C) in bits, or log10 two-sided Fisher p.  Host only."""
    s = ctypes.c_double(0.0)
    _check(lib().gepi_score_table(int(w), int(x), int(y), int(z), int(ffun), ctypes.byref(s)))
    return s.value


def _result_buffer(topmost, n_rows):
    total = 5 * (n_rows * (n_rows + 1) // 2)
    return np.zeros(max(1, min(int(topmost), total)), dtype=RESULT_DTYPE)


def _masks(mask_ctrl, mask_case, words):
    m0 = np.zeros(max(words, 1), np.uint64) if mask_ctrl is None else \
        np.ascontiguousarray(mask_ctrl, np.uint64)
    m1 = np.zeros(max(words, 1), np.uint64) if mask_case is None else \
        np.ascontiguousarray(mask_case, np.uint64)
    return m0, m1


def process_gpu_1(bit_data, valid_data, mask_ctrl, mask_case, n_samples, topmost=1000000,
                  ffun=GEPI_FFUNC_ENTROPY, gpu_num=0, count_path=0, max_grid_area=0):
    """Scans reader planes [R][W]; returns (results, stats dict)."""
    bits = np.ascontiguousarray(bit_data, np.uint64)
    valid = np.ascontiguousarray(valid_data, np.uint64)
    R = bits.shape[0]
    words = (int(n_samples) + 63) // 64
    m0, m1 = _masks(mask_ctrl, mask_case, words)
    par = make_params(topmost, ffun, gpu_num, count_path, max_grid_area)
    out = _result_buffer(topmost, R)
    n = ctypes.c_int64(0)
    st = ScanStats()
    b = bits if bits.size else np.zeros(1, np.uint64)
    v = valid if valid.size else np.zeros(1, np.uint64)
    _check(lib().gepi_process_gpu_1(_u64p(b), _u64p(v), _u64p(m0), _u64p(m1), R, int(n_samples),
                                    ctypes.byref(par), out.ctypes.data, ctypes.byref(n),
                                    ctypes.byref(st)))
    return out[:n.value].copy(), st.as_dict()


class Panel:
    """A genotype panel resident on one GPU across scans."""

    def __init__(self, handle, n_rows, n_samples):
        self._h = handle
        self.n_rows = n_rows
        self.n_samples = n_samples
        self.words = (n_samples + 63) // 64

    @classmethod
    def from_planes(cls, bit_data, valid_data, n_samples, device=0):
        bits = np.ascontiguousarray(bit_data, np.uint64)
        valid = np.ascontiguousarray(valid_data, np.uint64)
        h = ctypes.c_void_p()
        b = bits if bits.size else np.zeros(1, np.uint64)
        v = valid if valid.size else np.zeros(1, np.uint64)
        _check(lib().gepi_panel_create(_u64p(b), _u64p(v), bits.shape[0], int(n_samples),
                                       int(device), ctypes.byref(h)))
        return cls(h, bits.shape[0], int(n_samples))

    @classmethod
    def from_bed(cls, bed_payload, n_variants, n_samples, device=0):
        payload = np.ascontiguousarray(bed_payload, np.uint8)
        h = ctypes.c_void_p()
        buf = payload if payload.size else np.zeros(1, np.uint8)
        _check(lib().gepi_panel_create_bed(buf.ctypes.data, int(n_variants), int(n_samples),
                                           int(device), ctypes.byref(h)))
        return cls(h, 2 * int(n_variants), int(n_samples))

    def scan(self, mask_ctrl, mask_case, topmost=1000000, ffun=GEPI_FFUNC_ENTROPY, count_path=0,
             max_grid_area=0):
        m0, m1 = _masks(mask_ctrl, mask_case, self.words)
        par = make_params(topmost, ffun, 0, count_path, max_grid_area)
        out = _result_buffer(topmost, self.n_rows)
        n = ctypes.c_int64(0)
        st = ScanStats()
        _check(lib().gepi_scan(self._h, _u64p(m0), _u64p(m1), ctypes.byref(par),
                               out.ctypes.data, ctypes.byref(n), ctypes.byref(st)))
        return out[:n.value].copy(), st.as_dict()

    def scan_noalloc(self, mask_ctrl, mask_case, topmost=1000000, ffun=GEPI_FFUNC_ENTROPY,
                     count_path=0):
        """The same scan through gepi_scan_noalloc: no device allocation, all
        buffers in the library's fixed per-GPU scratch array (fp4 path)."""
        m0, m1 = _masks(mask_ctrl, mask_case, self.words)
        par = make_params(topmost, ffun, 0, count_path, 0)
        out = _result_buffer(topmost, self.n_rows)
        n = ctypes.c_int64(0)
        st = ScanStats()
        _check(lib().gepi_scan_noalloc(self._h, _u64p(m0), _u64p(m1), ctypes.byref(par),
                                       out.ctypes.data, ctypes.byref(n), ctypes.byref(st)))
        return out[:n.value].copy(), st.as_dict()

    def qa(self, max_miss_snps=0.05, max_miss_samples=0.05):
        """Missing calls per variant and per sample, and how many exceed the
        thresholds (gepi_panel_qa)."""
        n_var = self.n_rows // 2
        vm = np.zeros(max(n_var, 1), np.int32)
        sm = np.zeros(max(self.n_samples, 1), np.int32)
        bv, bs = ctypes.c_int64(0), ctypes.c_int64(0)
        i32p = ctypes.POINTER(ctypes.c_int32)
        _check(lib().gepi_panel_qa(self._h, float(max_miss_snps), float(max_miss_samples),
                                   vm.ctypes.data_as(i32p), sm.ctypes.data_as(i32p),
                                   ctypes.byref(bv), ctypes.byref(bs)))
        return vm[:n_var], sm[:self.n_samples], bv.value, bs.value

    def filter(self, max_miss_snps=0.05, max_miss_samples=0.05):
        """checkQuality(remove=true): (new Panel of the conforming variants,
        their original indices, mask of conforming samples)."""
        n_var = self.n_rows // 2
        kept = np.zeros(max(n_var, 1), np.int64)
        nk = ctypes.c_int64(0)
        keep = np.zeros(max(self.words, 1), np.uint64)
        h = ctypes.c_void_p()
        _check(lib().gepi_panel_filter(self._h, float(max_miss_snps), float(max_miss_samples),
                                       ctypes.byref(h), kept.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       ctypes.byref(nk), _u64p(keep)))
        return Panel(h, 2 * nk.value, self.n_samples), kept[:nk.value].copy(), keep

    def download(self):
        bits = np.zeros((self.n_rows, max(self.words, 1)), np.uint64)
        valid = np.zeros_like(bits)
        _check(lib().gepi_panel_download(self._h, _u64p(bits), _u64p(valid)))
        return bits[:, :self.words], valid[:, :self.words]

    def close(self):
        if self._h:
            _check(lib().gepi_panel_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """Batched scans on one device (gepi_engine_*): many scans advance together
    in rounds of three launches.  Each item's result equals Panel.scan's."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().gepi_engine_create(int(device), ctypes.byref(h)))
        self._h = h

    def scan_batch(self, items):
        """items: sequence of dicts {panel, mask_ctrl, mask_case, topmost, ffun}.
        Returns (list of (results, stats, status), batch stats dict); a failed
        item has status < 0 and no results."""
        n = len(items)
        arr = (BatchItem * max(n, 1))()
        keep = []
        for i, it in enumerate(items):
            panel = it["panel"]
            m0, m1 = _masks(it.get("mask_ctrl"), it.get("mask_case"), panel.words)
            topmost = int(it.get("topmost", 1000000))
            out = _result_buffer(topmost, panel.n_rows)
            keep.append((m0, m1, out))
            arr[i].panel = panel._h
            arr[i].mask_ctrl = _u64p(m0)
            arr[i].mask_case = _u64p(m1)
            arr[i].topmost = topmost
            arr[i].ffun = int(it.get("ffun", GEPI_FFUNC_ENTROPY))
            arr[i].tops = out.ctypes.data
        bs = BatchStats()
        _check(lib().gepi_engine_scan_batch(self._h, arr, n, ctypes.byref(bs)))
        res = []
        for i in range(n):
            out = keep[i][2]
            res.append((out[:arr[i].n_tops].copy(), arr[i].stats.as_dict(), int(arr[i].status)))
        return res, bs.as_dict()

    def autotune(self, repeats=2):
        """Runs the batch autotuner (gepi_engine_autotune) and returns its
        table: one dict per size class (max_ksteps, wg_per_cu, tile,
        band_evals, ms_best, ms_default), ascending max_ksteps; tile 0 is the
        64 x 32-variant count job, 1 the 32 x 32 half tile, 2 wave tiles (two
        64 x 32 tiles as four 32 x 32 wave tiles, the kernel of count path 3).
        tile schedules entropy scans; Fisher scans run wave tiles in every
        class (not in captured-graph rounds)."""
        out = (TuneClass * 16)()
        n = ctypes.c_int32()
        _check(lib().gepi_engine_autotune(self._h, int(repeats), out, 16, ctypes.byref(n)))
        return [out[i].as_dict() for i in range(min(n.value, 16))]

    def set_tuning(self, classes):
        """Installs a table of size classes (dicts as autotune returns;
        an empty list restores the defaults)."""
        arr = (TuneClass * max(len(classes), 1))()
        for i, c in enumerate(classes):
            arr[i].max_ksteps = int(c["max_ksteps"])
            arr[i].wg_per_cu = int(c["wg_per_cu"])
            arr[i].tile = int(c.get("tile", 0))
            arr[i].band_evals = float(c["band_evals"])
            arr[i].ms_best = float(c.get("ms_best", 0.0))
            arr[i].ms_default = float(c.get("ms_default", 0.0))
        _check(lib().gepi_engine_set_tuning(self._h, arr, len(classes)))

    def close(self):
        if self._h:
            _check(lib().gepi_engine_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop (gepi_server_*): resident panels keyed by content hash,
    requests admitted between rounds of the batched engine by the server's
    worker thread.  Thread safe."""

    def __init__(self, device=0, hbm_budget=0, round_wg_budget=0, autotune=False):
        par = ServerParams(int(device), int(bool(autotune)), int(hbm_budget), int(round_wg_budget))
        h = ctypes.c_void_p()
        _check(lib().gepi_server_create(ctypes.byref(par), ctypes.byref(h)))
        self._h = h
        self._words = {}

    def add_panel_bed(self, bed_payload, n_variants, n_samples):
        payload = np.ascontiguousarray(bed_payload, np.uint8)
        buf = payload if payload.size else np.zeros(1, np.uint8)
        pid = ctypes.c_uint64()
        _check(lib().gepi_server_add_panel_bed(self._h, buf.ctypes.data, int(n_variants),
                                               int(n_samples), ctypes.byref(pid)))
        self._words[pid.value] = ((int(n_samples) + 63) // 64, 2 * int(n_variants))
        return pid.value

    def release_panel(self, panel_id):
        _check(lib().gepi_server_release_panel(self._h, ctypes.c_uint64(panel_id)))

    def submit(self, panel_id, mask_ctrl, mask_case, topmost=1000, ffun=GEPI_FFUNC_ENTROPY):
        # an unknown id is the library's to reject (GEPI_ERR_ARG)
        words, rows = self._words.get(panel_id, (1, 0))
        m0, m1 = _masks(mask_ctrl, mask_case, words)
        t = ctypes.c_uint64()
        _check(lib().gepi_server_submit(self._h, ctypes.c_uint64(panel_id), _u64p(m0), _u64p(m1),
                                        int(topmost), int(ffun), ctypes.byref(t)))
        return (t.value, int(topmost), rows)

    def done(self, ticket):
        d = ctypes.c_int32(0)
        _check(lib().gepi_server_done(self._h, ctypes.c_uint64(ticket[0]), ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        """Blocks until the request is done; returns (results, stats dict, status)."""
        t, topmost, rows = ticket
        out = _result_buffer(topmost, rows)
        n = ctypes.c_int64(0)
        st = ScanStats()
        rc = lib().gepi_server_wait(self._h, ctypes.c_uint64(t), out.ctypes.data, len(out),
                                    ctypes.byref(n), ctypes.byref(st))
        return out[:n.value].copy(), st.as_dict(), int(rc)

    def stats(self):
        st = ServerStats()
        _check(lib().gepi_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self._h:
            _check(lib().gepi_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlinkData:
    """Reader output (BinPlinkReader): planes, class masks and SNP/sample info."""

    def __init__(self, root):
        pp = ctypes.POINTER(PlinkStruct)()
        _check(lib().gepi_plink_open(os.fsencode(root), ctypes.byref(pp)))
        self._p = pp
        p = pp.contents
        R, W, n = p.n_rows, p.words, p.n_samples
        self.n_rows, self.n_samples, self.words = R, n, W
        self.bit_data = (np.ctypeslib.as_array(p.bit_data, shape=(R * W,)).reshape(R, W).copy()
                         if R * W else np.zeros((R, W), np.uint64))
        self.valid_data = (np.ctypeslib.as_array(p.valid_data, shape=(R * W,)).reshape(R, W).copy()
                           if R * W else np.zeros((R, W), np.uint64))
        self.mask_ctrl = np.ctypeslib.as_array(p.mask_ctrl, shape=(W,)).copy() if p.mask_ctrl else None
        self.mask_case = np.ctypeslib.as_array(p.mask_case, shape=(W,)).copy() if p.mask_case else None
        nbytes = (R // 2) * ((n + 3) // 4)
        self.bed_payload = (np.ctypeslib.as_array(p.bed_payload, shape=(nbytes,)).copy()
                            if nbytes else np.zeros(0, np.uint8))
        self.snp_id = [p.snp_id[i].decode() for i in range(R)]
        self.allele = [p.allele[i].decode() for i in range(R)]
        self.chrom = [p.chr[i] for i in range(R)]
        self.pos = [p.pos[i] for i in range(R)]
        self.pheno = [p.pheno[i] for i in range(n)]
        self.sample_name = [p.sample_name[i].decode() for i in range(n)]
        self.total_missing = p.total_missing

    def check_quality(self, max_miss_snps=0.05, max_miss_samples=0.05):
        a, b = ctypes.c_int32(0), ctypes.c_int32(0)
        _check(lib().gepi_plink_check_quality(self._p, max_miss_snps, max_miss_samples,
                                              ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        if self._p:
            lib().gepi_plink_close(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_plink(root):
    return PlinkData(root)


def fisher_pval(n11, n12, n21, n22):
    return lib().gepi_fisher_pval(int(n11), int(n12), int(n21), int(n22))
