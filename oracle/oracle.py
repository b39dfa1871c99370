"""ctypes wrapper around oracle/liboracle.so, plus pure-Python brute force.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  See oracle.h for what is
restated from which reference file.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GEPI_ORACLE_LIB names another build of the same sources (for instance one
# compiled for the local CPU to run a very large gold scan).
LIB_PATH = os.environ.get("GEPI_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_FISHER_PATH = os.path.join(HERE, "_ref", "libfisher_ref.so")

RESULT_DTYPE = np.dtype([("sA", "<i8"), ("sB", "<i8"), ("w", "<i4"), ("x", "<i4"),
                         ("y", "<i4"), ("z", "<i4"), ("obj", "<f8"), ("fun", "<i4"),
                         ("reserved", "<i4")])
FUNC_NAMES = ("AND", "NAND", "ANDN", "NANDN", "XOR")


class _Plink(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int64), ("n_samples", ctypes.c_int32),
                ("words", ctypes.c_int32),
                ("bits", ctypes.POINTER(ctypes.c_uint64)),
                ("valid", ctypes.POINTER(ctypes.c_uint64)),
                ("mask1", ctypes.POINTER(ctypes.c_uint64)),
                ("mask2", ctypes.POINTER(ctypes.c_uint64)),
                ("pheno", ctypes.POINTER(ctypes.c_int32)),
                ("sample_name", ctypes.POINTER(ctypes.c_char_p)),
                ("snp_id", ctypes.POINTER(ctypes.c_char_p)),
                ("allele", ctypes.POINTER(ctypes.c_char)),
                ("chr", ctypes.POINTER(ctypes.c_int32)),
                ("pos", ctypes.POINTER(ctypes.c_int32)),
                ("total_missing", ctypes.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        l = ctypes.CDLL(LIB_PATH)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        l.orc_plink_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(_Plink))]
        l.orc_plink_read.restype = ctypes.c_int
        l.orc_plink_free.argtypes = [ctypes.POINTER(_Plink)]
        l.orc_last_error.restype = ctypes.c_char_p
        l.orc_entropy_exact.argtypes = [ctypes.c_int] * 4
        l.orc_entropy_exact.restype = ctypes.c_double
        l.orc_entropy_approx.argtypes = [ctypes.c_int] * 4
        l.orc_entropy_approx.restype = ctypes.c_double
        l.orc_entropy_ref_f32.argtypes = [ctypes.c_int] * 4
        l.orc_entropy_ref_f32.restype = ctypes.c_float
        l.orc_fisher_ref.argtypes = [ctypes.c_int] * 4
        l.orc_fisher_ref.restype = ctypes.c_float
        l.orc_fisher_rank.argtypes = [ctypes.c_int] * 4
        l.orc_fisher_rank.restype = ctypes.c_double
        l.orc_log2.argtypes = [ctypes.c_double]
        l.orc_log2.restype = ctypes.c_double
        l.orc_gold_scan.argtypes = [u64p, u64p, u64p, u64p, ctypes.c_int64, ctypes.c_int32,
                                    ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.POINTER(ctypes.c_int64)]
        l.orc_gold_scan.restype = ctypes.c_int
        l.orc_gold_scan_rows.argtypes = [u64p, u64p, u64p, u64p, ctypes.c_int64,
                                         ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        l.orc_gold_scan_rows.restype = ctypes.c_int
        l.orc_gold_scan_ref_f32.argtypes = [u64p, u64p, u64p, u64p, ctypes.c_int64,
                                            ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_double)]
        l.orc_gold_scan_ref_f32.restype = ctypes.c_int
        l.orc_all_counts.argtypes = [u64p, u64p, u64p, u64p, ctypes.c_int64, ctypes.c_int32,
                                     ctypes.c_void_p]
        l.orc_all_counts.restype = ctypes.c_int
        _lib = l
    return _lib


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class Panel:
    """Allele-row bit planes in the reference host layout [R][W] (PlinkReader.h:142-149)."""

    def __init__(self, bits, valid, mask_ctrl, mask_case, n_samples, snp_id=None,
                 allele=None, chrom=None, pos=None, pheno=None, sample_name=None):
        self.bits = np.ascontiguousarray(bits, dtype=np.uint64)
        self.valid = np.ascontiguousarray(valid, dtype=np.uint64)
        self.mask_ctrl = None if mask_ctrl is None else np.ascontiguousarray(mask_ctrl, np.uint64)
        self.mask_case = None if mask_case is None else np.ascontiguousarray(mask_case, np.uint64)
        self.n_samples = int(n_samples)
        self.snp_id, self.allele, self.chrom, self.pos = snp_id, allele, chrom, pos
        self.pheno, self.sample_name = pheno, sample_name

    @property
    def n_rows(self):
        return self.bits.shape[0]

    @property
    def words(self):
        return self.bits.shape[1]


def read_plink(root):
    l = lib()
    pp = ctypes.POINTER(_Plink)()
    if l.orc_plink_read(root.encode(), ctypes.byref(pp)) != 0:
        raise ValueError(l.orc_last_error().decode())
    p = pp.contents
    R, W, n = p.n_rows, p.words, p.n_samples
    bits = np.ctypeslib.as_array(p.bits, shape=(R * W,)).copy().reshape(R, W) if R * W else \
        np.zeros((R, W), np.uint64)
    valid = np.ctypeslib.as_array(p.valid, shape=(R * W,)).copy().reshape(R, W) if R * W else \
        np.zeros((R, W), np.uint64)
    m1 = np.ctypeslib.as_array(p.mask1, shape=(W,)).copy() if p.mask1 else None
    m2 = np.ctypeslib.as_array(p.mask2, shape=(W,)).copy() if p.mask2 else None
    panel = Panel(bits, valid, m1, m2, n,
                  snp_id=[p.snp_id[i].decode() for i in range(R)],
                  allele=[p.allele[i].decode() for i in range(R)],
                  chrom=[p.chr[i] for i in range(R)], pos=[p.pos[i] for i in range(R)],
                  pheno=[p.pheno[i] for i in range(n)],
                  sample_name=[p.sample_name[i].decode() for i in range(n)])
    panel.total_missing = p.total_missing
    l.orc_plink_free(pp)
    return panel


def gold_scan(panel, topk, ffun=0, nthreads=0):
    """Restated process_gold_1; returns a RESULT_DTYPE array sorted by canonical key."""
    l = lib()
    R, W = panel.n_rows, panel.words
    total = 5 * (R * (R + 1) // 2)
    out = np.zeros(max(1, min(topk, total)), dtype=RESULT_DTYPE)
    n = ctypes.c_int64(0)
    m0 = panel.mask_ctrl if panel.mask_ctrl is not None else np.zeros(max(W, 1), np.uint64)
    m1 = panel.mask_case if panel.mask_case is not None else np.zeros(max(W, 1), np.uint64)
    bits = panel.bits if panel.bits.size else np.zeros(1, np.uint64)
    valid = panel.valid if panel.valid.size else np.zeros(1, np.uint64)
    rc = l.orc_gold_scan(_u64p(bits), _u64p(valid), _u64p(m0), _u64p(m1), R, W, topk, ffun,
                         nthreads, out.ctypes.data, ctypes.byref(n))
    if rc != 0:
        raise ValueError("orc_gold_scan failed")
    return out[:n.value].copy()


def gold_scan_rows(panel, j0, j1, topk, ffun=0, nthreads=0):
    """The gold scan over rows j in [j0, j1) (every k <= j): the top-k of that
    part of the scan, sorted by canonical key."""
    l = lib()
    R, W = panel.n_rows, panel.words
    total = 5 * (j1 * (j1 + 1) // 2 - j0 * (j0 + 1) // 2)
    out = np.zeros(max(1, min(topk, total)), dtype=RESULT_DTYPE)
    n = ctypes.c_int64(0)
    m0 = panel.mask_ctrl if panel.mask_ctrl is not None else np.zeros(max(W, 1), np.uint64)
    m1 = panel.mask_case if panel.mask_case is not None else np.zeros(max(W, 1), np.uint64)
    rc = l.orc_gold_scan_rows(_u64p(panel.bits), _u64p(panel.valid), _u64p(m0), _u64p(m1), R, W,
                              j0, j1, topk, ffun, nthreads, out.ctypes.data, ctypes.byref(n))
    if rc != 0:
        raise ValueError("orc_gold_scan_rows failed")
    return out[:n.value].copy()


def gold_scan_ref_f32(panel, topk, j_end=None):
    """The reference's single-thread float32 gold scan over rows j < j_end.
    Returns (results with obj = the float32 score, seconds)."""
    l = lib()
    R, W = panel.n_rows, panel.words
    j_end = R if j_end is None else int(j_end)
    out = np.zeros(max(1, topk), dtype=RESULT_DTYPE)
    n = ctypes.c_int64(0)
    sec = ctypes.c_double(0.0)
    m0 = panel.mask_ctrl if panel.mask_ctrl is not None else np.zeros(max(W, 1), np.uint64)
    m1 = panel.mask_case if panel.mask_case is not None else np.zeros(max(W, 1), np.uint64)
    rc = l.orc_gold_scan_ref_f32(_u64p(panel.bits), _u64p(panel.valid), _u64p(m0), _u64p(m1), R,
                                 W, topk, j_end, out.ctypes.data, ctypes.byref(n),
                                 ctypes.byref(sec))
    if rc != 0:
        raise ValueError("orc_gold_scan_ref_f32 failed")
    return out[:n.value].copy(), sec.value


def all_counts(panel):
    """Counts [P, 5, 4] in gold scan order (j, k <= j, fun)."""
    l = lib()
    R, W = panel.n_rows, panel.words
    P = R * (R + 1) // 2
    out = np.zeros((max(P, 1), 5, 4), dtype=np.int32)
    m0 = panel.mask_ctrl if panel.mask_ctrl is not None else np.zeros(max(W, 1), np.uint64)
    m1 = panel.mask_case if panel.mask_case is not None else np.zeros(max(W, 1), np.uint64)
    l.orc_all_counts(_u64p(panel.bits), _u64p(panel.valid), _u64p(m0), _u64p(m1), R, W,
                     out.ctypes.data)
    return out[:P]


def entropy_exact(w, x, y, z):
    return lib().orc_entropy_exact(w, x, y, z)


def entropy_ref_f32(w, x, y, z):
    return lib().orc_entropy_ref_f32(w, x, y, z)


def fisher_ref(n11, n12, n21, n22):
    return lib().orc_fisher_ref(n11, n12, n21, n22)


def fisher_rank(w, x, y, z):
    return lib().orc_fisher_rank(w, x, y, z)


# ---------------------------------------------------------------------------
# Pure-Python brute force from per-sample genotype codes (small inputs only).

def planes_from_codes(codes):
    """codes: [N variants][n samples] PLINK 2-bit codes (0 hom A1, 1 missing, 2 het,
    3 hom A2).  Returns (bits, valid) [2N][W] following BinPlinkReader.h:171-201."""
    codes = np.asarray(codes)
    N, n = codes.shape
    W = (n + 63) // 64
    bits = np.zeros((2 * N, W), np.uint64)
    valid = np.zeros((2 * N, W), np.uint64)
    for i in range(N):
        for s in range(n):
            c = int(codes[i, s])
            bit = np.uint64(1) << np.uint64(s % 64)
            if c == 1:
                continue
            valid[2 * i, s // 64] |= bit
            valid[2 * i + 1, s // 64] |= bit
            if c in (0, 2):
                bits[2 * i, s // 64] |= bit
            if c in (2, 3):
                bits[2 * i + 1, s // 64] |= bit
    return bits, valid


def brute_counts(codes, cls):
    """Per-sample loop: for every row pair (j, k<=j) and function, the counts
    (w, x, y, z) straight from the genotype codes.  cls[s] in {0 ctrl, 1 case, -1}."""
    codes = np.asarray(codes)
    N, n = codes.shape
    carries = np.zeros((2 * N, n), bool)
    valid = np.zeros((2 * N, n), bool)
    for i in range(N):
        c = codes[i]
        valid[2 * i] = valid[2 * i + 1] = c != 1
        carries[2 * i] = (c == 0) | (c == 2)
        carries[2 * i + 1] = (c == 2) | (c == 3)
    cls = np.asarray(cls)
    funcs = (lambda a, b: a & b, lambda a, b: ~a & b, lambda a, b: a & ~b,
             lambda a, b: ~(a | b), lambda a, b: a ^ b)
    R = 2 * N
    out = np.zeros((R * (R + 1) // 2, 5, 4), np.int64)
    e = 0
    for j in range(R):
        for k in range(j + 1):
            both = valid[j] & valid[k]
            for f, fn in enumerate(funcs):
                v = fn(carries[j], carries[k]) & both
                out[e, f] = (np.sum(v & (cls == 0)), np.sum(v & (cls == 1)),
                             np.sum(both & (cls == 0)), np.sum(both & (cls == 1)))
            e += 1
    return out


def masks_from_classes(cls):
    cls = np.asarray(cls)
    n = len(cls)
    W = (n + 63) // 64
    m0 = np.zeros(W, np.uint64)
    m1 = np.zeros(W, np.uint64)
    for s in range(n):
        bit = np.uint64(1) << np.uint64(s % 64)
        if cls[s] == 0:
            m0[s // 64] |= bit
        elif cls[s] == 1:
            m1[s // 64] |= bit
    return m0, m1


def random_codes(rng, N, n, miss=0.0):
    maf = rng.uniform(0.05, 0.5, size=N)
    g = rng.binomial(2, maf[:, None], size=(N, n))
    codes = np.where(g == 0, 0, np.where(g == 1, 2, 3))
    if miss > 0:
        codes = np.where(rng.random((N, n)) < miss, 1, codes)
    return codes.astype(np.int64)


def report_lines(panel, tops):
    """The print_misc.h:56-78 report as a list of lines (header first)."""
    lines = ["", "Rank:\tidxA[id allele chr:pos]\tFunc\tidxB[id allele chr:pos]\tObj\tFPv\t"
                 "(c00,c01,c10,c11)"]
    for i, r in enumerate(tops):
        a, b = int(r["sA"]), int(r["sB"])
        w, x, y, z = int(r["w"]), int(r["x"]), int(r["y"]), int(r["z"])
        pv = fisher_ref(w, y - w, x, z - x)
        lines.append("%d:\t%d[%s %s %d:%d]\t%s\t%d[%s %s %d:%d]\t%s\t%s\t(%d,%d,%d,%d)" % (
            i, a, panel.snp_id[a], panel.allele[a], panel.chrom[a], panel.pos[a],
            FUNC_NAMES[int(r["fun"])], b, panel.snp_id[b], panel.allele[b], panel.chrom[b],
            panel.pos[b], _g(float(np.float32(r["obj"]))), _g(pv), w, y - w, x, z - x))
    return lines


def _g(v):
    return "%g" % v
