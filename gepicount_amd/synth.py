"""Seeded synthetic PLINK panels for the scan workloads.

Genotypes are Hardy-Weinberg Binomial(2, MAF) counts of the minor allele,
which is the first .bim allele (A1): 2 copies -> code 0 (hom A1), 1 -> code 2
(het), 0 -> code 3 (hom A2), missing -> code 1.  Samples get phenotype 1
(control) or 2 (case).  Workloads (variants x samples):

  small   1,024 x 256 (128/128), MAF ~ U[0.05, 0.5], no missing calls
  medium  20,000 x 2,048 (1,024/1,024), MAF ~ U[0.01, 0.5], 1% missing,
          5 planted XOR pairs with case/control odds ratio 2.0
  large   100,000 x 8,192 (2,458 cases / 5,734 controls), MAF ~ Beta(0.5, 2)
          clipped to [0.01, 0.5], 2% missing

and the mix of uneven scans (mix_specs): variants ~ logU[500, 50,000],
samples ~ logU[200, 5,000], case fraction ~ U[0.2, 0.8], MAF ~ U[0.01, 0.5],
1% missing, top-100 each.
"""
import os

import numpy as np

WORKLOADS = {
    "small": dict(n_var=1024, n_samples=256, n_case=128, maf=("uniform", 0.05, 0.5), miss=0.0,
                  planted=0, seed=1001),
    "medium": dict(n_var=20000, n_samples=2048, n_case=1024, maf=("uniform", 0.01, 0.5), miss=0.01,
                   planted=5, seed=1002),
    "large": dict(n_var=100000, n_samples=8192, n_case=2458, maf=("beta", 0.5, 2.0), miss=0.02,
                  planted=0, seed=1003),
}


def _maf(rng, spec, n):
    kind, a, b = spec
    if kind == "uniform":
        return rng.uniform(a, b, size=n)
    return np.clip(rng.beta(a, b, size=n), 0.01, 0.5)


def pack_codes(codes):
    """[V][n] PLINK codes (uint8 0..3) -> SNP-major .bed payload [V][ceil(n/4)]."""
    codes = np.asarray(codes, dtype=np.uint8)
    V, n = codes.shape
    n4 = (n + 3) // 4
    pad = np.zeros((V, n4 * 4), np.uint8)
    pad[:, :n] = codes
    q = pad.reshape(V, n4, 4)
    return (q[:, :, 0] | (q[:, :, 1] << 2) | (q[:, :, 2] << 4) | (q[:, :, 3] << 6)).astype(np.uint8)


def _codes_chunk(rng, maf, n, miss):
    u = rng.random((len(maf), n), dtype=np.float32)
    v = rng.random((len(maf), n), dtype=np.float32)
    m = maf.astype(np.float32)[:, None]
    g = (u < m).astype(np.uint8) + (v < m).astype(np.uint8)
    codes = np.where(g == 2, 0, np.where(g == 1, 2, 3)).astype(np.uint8)
    if miss > 0:
        codes[rng.random((len(maf), n), dtype=np.float32) < miss] = 1
    return codes


def _plant_xor(rng, codes, maf, cls, a, b, odds=2.0):
    """Redraws variants a and b so that carrier(a) XOR carrier(b) has the given
    odds ratio between cases and controls (importance resampling of cases)."""
    n = codes.shape[1]
    ga = rng.binomial(2, maf[a], size=n)
    gb = rng.binomial(2, maf[b], size=n)
    case = np.flatnonzero(cls == 1)
    for s in case:
        while True:
            xa, xb = rng.binomial(2, maf[a]), rng.binomial(2, maf[b])
            xor = (xa > 0) != (xb > 0)
            if rng.random() < (1.0 if xor else 1.0 / odds):
                ga[s], gb[s] = xa, xb
                break
    for v, g in ((a, ga), (b, gb)):
        c = np.where(g == 2, 0, np.where(g == 1, 2, 3)).astype(np.uint8)
        keep_missing = codes[v] == 1
        c[keep_missing] = 1
        codes[v] = c


def mix_specs(n_scans=64, seed=1004):
    """Specs of the uneven-scan mix: list of dicts usable as make_panel(spec=...)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_scans):
        V = int(round(np.exp(rng.uniform(np.log(500), np.log(50000)))))
        n = int(round(np.exp(rng.uniform(np.log(200), np.log(5000)))))
        n_case = int(round(rng.uniform(0.2, 0.8) * n))
        out.append(dict(n_var=V, n_samples=n, n_case=n_case, maf=("uniform", 0.01, 0.5),
                        miss=0.01, planted=0, seed=seed * 1000 + i, topk=100))
    return out


def pheno_for(n, n_case, seed):
    """A seeded phenotype vector with n_case cases (2) among n samples (1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cls = np.zeros(n, np.int64)
    cls[rng.permutation(n)[:n_case]] = 1
    return np.where(cls == 1, 2, 1)


def make_panel(workload=None, n_var=None, seed=None, chunk=4096, spec=None):
    """Returns dict(payload [V][ceil(n/4)] uint8, pheno [n] int, n_var, n_samples,
    planted [(a, b)], maf).  `spec` (a dict like WORKLOADS' entries) replaces
    the named workload."""
    spec = dict(WORKLOADS[workload] if spec is None else spec)
    if n_var is not None:
        spec["n_var"] = int(n_var)
    rng = np.random.Generator(np.random.PCG64(spec["seed"] if seed is None else seed))
    V, n, n_case = spec["n_var"], spec["n_samples"], spec["n_case"]
    maf = _maf(rng, spec["maf"], V)
    cls = np.zeros(n, np.int64)
    cls[rng.permutation(n)[:n_case]] = 1
    pheno = np.where(cls == 1, 2, 1)
    n4 = (n + 3) // 4
    payload = np.zeros((V, n4), np.uint8)
    planted = []
    for v0 in range(0, V, chunk):
        v1 = min(V, v0 + chunk)
        codes = _codes_chunk(rng, maf[v0:v1], n, spec["miss"])
        payload[v0:v1] = pack_codes(codes)
    if spec["planted"]:
        prng = np.random.Generator(np.random.PCG64(spec["seed"] + 77))
        idx = prng.choice(V, size=2 * spec["planted"], replace=False)
        for i in range(spec["planted"]):
            a, b = int(idx[2 * i]), int(idx[2 * i + 1])
            codes = unpack_codes(payload[[a, b]], n)
            sub = codes.copy()
            _plant_xor(prng, sub, maf[[a, b]], cls, 0, 1)
            payload[[a, b]] = pack_codes(sub)
            planted.append((a, b))
    return dict(payload=payload, pheno=pheno, n_var=V, n_samples=n, planted=planted, maf=maf)


def unpack_codes(payload, n):
    payload = np.asarray(payload, np.uint8)
    V = payload.shape[0]
    q = np.stack([(payload >> (2 * i)) & 3 for i in range(4)], axis=-1).reshape(V, -1)
    return q[:, :n].copy()


def masks_from_pheno(pheno):
    """Class masks (phenotype 1 -> control, 2 -> case) as [ceil(n/64)] uint64."""
    pheno = np.asarray(pheno)
    n = len(pheno)
    W = (n + 63) // 64
    bits = np.zeros(W * 64, np.uint8)
    out = []
    for code in (1, 2):
        bits[:] = 0
        bits[:n] = pheno == code
        out.append(np.packbits(bits.reshape(W, 64)[:, ::-1], axis=1).view(">u8").ravel()
                   .astype(np.uint64))
    return out[0], out[1]


def write_plink(root, panel):
    """Writes <root>.bed/.bim/.fam for a panel from make_panel."""
    os.makedirs(os.path.dirname(os.path.abspath(root)), exist_ok=True)
    V, n = panel["n_var"], panel["n_samples"]
    with open(root + ".bed", "wb") as f:
        f.write(bytes([0x6C, 0x1B, 0x01]))
        f.write(np.ascontiguousarray(panel["payload"]).tobytes())
    with open(root + ".bim", "w") as f:
        for i in range(V):
            f.write("1\trs%d\t0\t%d\tA\tG\n" % (i + 1, 1000 + 10 * i))
    with open(root + ".fam", "w") as f:
        for s in range(n):
            f.write("FID%d IID%d 0 0 1 %d\n" % (s + 1, s + 1, int(panel["pheno"][s])))
