/*
 * gepi_stage.h -- panel staging kernels (replace the host transpose and
 * uploads of src/kernels/setup1.h:72-102 and the constant-memory class masks
 * of kernel1.cu:329-337).
 *
 * k_decode_bed     SNP-major .bed bytes -> [n_var][3][W64] planes on device
 *                  (BinPlinkReader.h:171-201: code 0 -> allele 1 + valid,
 *                  2 -> both + valid, 3 -> allele 2 + valid, 1 -> missing).
 * k_gather_class   per scan: samples reordered class-contiguously
 *                  (class 0 then class 1, each padded to 128 bits) into the
 *                  popcount layout [n_var_pad][3][wc] dwords.
 * k_transpose_k    popcount layout -> MFMA layout [n_ks][n_var_pad][3] of
 *                  64-sample words, so that one K step of a tile is contiguous.
 * k_qa_variants    missing calls per variant (one wave per variant)
 * k_qa_samples     missing calls per sample (a workgroup per 256 samples and
 *                  a chunk of variants; partial counts added atomically)
 * k_gather_variants  copies the planes of selected variants to a new panel
 * k_class_f4       panel -> fp4 MFMA layout directly: per plane row, the
 *                  samples of each class are compressed (stable) out of the
 *                  file order with mask-driven bit compression, cut into
 *                  128-sample K steps and written four K steps at a time as
 *                  64-byte super-step rows (class 0, then class 1 from the
 *                  next super-step), with each row's set bits in the fourth
 *                  K steps; or three K steps at a time in the grouped
 *                  layout of the wave-tile count kernel.
 * The QA kernels replace the loops of checkQuality (src/reader/
 * PlinkReader.h:237-335), which count missing calls on the host.
 */
#ifndef GEPI_STAGE_H
#define GEPI_STAGE_H

#include "gepi_device.h"

namespace gepi {

static __global__ void k_decode_bed(const uint8_t *__restrict__ bed, int64_t n_var, int32_t n_samples,
                             int32_t w64, int64_t bps, uint64_t *__restrict__ nat) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_var * w64) return;
    int64_t v = gid / w64;
    int32_t wd = (int32_t)(gid % w64);
    const uint8_t *src = bed + v * bps;
    uint64_t a1 = 0, a2 = 0, vv = 0;
    int32_t s0 = wd * 64;
    int32_t s1 = s0 + 64 < n_samples ? s0 + 64 : n_samples;
    for (int32_t s = s0; s < s1; s += 4) {
        uint32_t byte = src[s >> 2];
        for (int32_t q = 0; q < 4 && s + q < s1; q++) {
            uint32_t code = (byte >> (2 * q)) & 3u;
            uint64_t bit = 1ull << ((s + q) & 63);
            if (code == 1u) continue;
            vv |= bit;
            if (code != 3u) a1 |= bit;
            if (code != 0u) a2 |= bit;
        }
    }
    uint64_t *dst = nat + (v * 3) * w64 + wd;
    dst[0] = a1;
    dst[w64] = a2;
    dst[2 * w64] = vv;
}

/* out[v][p][d] bit b = panel bit perm[32*d + b] (0 for perm < 0 or v >= n_var). */
static __global__ void __launch_bounds__(256) k_gather_class(const uint64_t *__restrict__ nat,
                                                       int32_t n_var, int32_t w64,
                                                       const int32_t *__restrict__ perm,
                                                       int32_t wc, int32_t wused,
                                                       uint32_t *__restrict__ out) {
    extern __shared__ uint64_t srow[];
    const int32_t v = blockIdx.x;
    const bool real = v < n_var;
    for (int32_t i = threadIdx.x; i < 3 * w64; i += blockDim.x)
        srow[i] = real ? nat[(int64_t)v * 3 * w64 + i] : 0ull;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < 3 * wc; i += blockDim.x) {
        int32_t p = i / wc, d = i - p * wc;
        uint32_t word = 0;
        if (real && d < wused) {
            const int32_t *pp = perm + 32 * d;
            const uint64_t *row = srow + p * w64;
#pragma unroll 8
            for (int32_t b = 0; b < 32; b++) {
                int32_t s = pp[b];
                if (s >= 0) word |= (uint32_t)((row[s >> 6] >> (s & 63)) & 1ull) << b;
            }
        }
        out[((int64_t)v * 3 + p) * wc + d] = word;
    }
}

/* Tile of 32 variants x 32 K steps per workgroup. */
static __global__ void __launch_bounds__(256) k_transpose_k(const uint32_t *__restrict__ cls,
                                                      int32_t n_var_pad, int32_t wc, int32_t n_ks,
                                                      uint64_t *__restrict__ out) {
    __shared__ uint64_t tile[96][33];
    const int32_t v0 = blockIdx.x * 32, s0 = blockIdx.y * 32;
    for (int32_t idx = threadIdx.x; idx < 96 * 32; idx += blockDim.x) {
        const int32_t vp = idx >> 5, k = idx & 31;
        const int32_t s = s0 + k;
        const int32_t v = v0 + vp / 3, p = vp % 3;
        uint64_t val = 0;
        if (s < n_ks)
            val = *reinterpret_cast<const uint64_t *>(cls + ((int64_t)v * 3 + p) * wc + 2 * s);
        tile[vp][k] = val;
    }
    __syncthreads();
    for (int32_t idx = threadIdx.x; idx < 96 * 32; idx += blockDim.x) {
        const int32_t k = idx / 96, vp = idx - 96 * k;
        const int32_t s = s0 + k;
        if (s < n_ks) out[((int64_t)s * n_var_pad + v0) * 3 + vp] = tile[vp][k];
    }
}

/* Compression of one 32-bit sample word for one class: x & m moved to the
 * low popc(m) bits, order kept, by five conditional shifts whose move masks
 * depend on m only (Hacker's Delight, "compress").  Built on the host once
 * per scan; the same for every plane row.                                   */
struct CompressWord {
    uint32_t m;     /* class mask word                                      */
    uint32_t mv[5]; /* move masks of the stages shifting by 1, 2, 4, 8, 16 */
    uint32_t cnt;   /* popc(m)                                              */
    uint32_t pad;
};

__host__ __device__ inline CompressWord make_compress_word(uint32_t m) {
    CompressWord c;
    c.m = m;
    c.cnt = 0;
    for (uint32_t t = m; t; t &= t - 1) c.cnt++;
    uint32_t mk = ~m << 1; /* zeros to the right of each position */
    for (int i = 0; i < 5; i++) {
        uint32_t mp = mk ^ (mk << 1); /* parallel suffix (xor) */
        mp ^= mp << 2;
        mp ^= mp << 4;
        mp ^= mp << 8;
        mp ^= mp << 16;
        const uint32_t mv = mp & m;
        c.mv[i] = mv;
        m = (m ^ mv) | (mv >> (1u << i));
        mk &= ~mp;
    }
    c.pad = 0;
    return c;
}

__device__ __forceinline__ uint32_t compress_word(uint32_t x, const CompressWord &c) {
    x &= c.m;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const uint32_t t = x & c.mv[i];
        x = (x ^ t) | (t >> (1u << i));
    }
    return x;
}

/* Bit n of the low byte of x moved to bit 4n (three shift-or-mask rounds). */
__device__ __forceinline__ uint32_t spread8(uint32_t x) {
    x &= 0xFFu;
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x;
}

/* One 64-byte super-step row of the fp4 layout from the row's four
 * 128-sample K steps k0..k3 (32-bit word g of a K step holds its samples
 * 32g .. 32g+31): dword d of 16-byte group g holds, at bit 4n + p, sample
 * 32g + 8d + n of K step p (bit 3, the e2m1 sign bit, for K step 3).      */
__device__ __forceinline__ uint4 super_group(uint32_t a, uint32_t b, uint32_t c, uint32_t e) {
    uint4 o;
    o.x = spread8(a) | (spread8(b) << 1) | (spread8(c) << 2) | (spread8(e) << 3);
    o.y = spread8(a >> 8) | (spread8(b >> 8) << 1) | (spread8(c >> 8) << 2) | (spread8(e >> 8) << 3);
    o.z = spread8(a >> 16) | (spread8(b >> 16) << 1) | (spread8(c >> 16) << 2) |
          (spread8(e >> 16) << 3);
    o.w = spread8(a >> 24) | (spread8(b >> 24) << 1) | (spread8(c >> 24) << 2) |
          (spread8(e >> 24) << 3);
    return o;
}

/* A stream of bits appended in order, cut into 128-sample K steps, and
 * written S (3 or 4) K steps at a time as super-step rows
 * out[(u * rows + r) * 4 + g]; with S = 4, pc3 counts the set bits of every
 * fourth K step (the count kernel's correction of that step).  GRP: the
 * grouped layout of the wave-tile kernel instead, uint4 g of row r = 3v + p
 * at out[(((u * NG + v / 32) * 3 + p) * 4 + g) * 32 + v % 32] (NG = rows /
 * 96 groups of 32 variants), so that the 32 variants of a group, one plane
 * and one 16-byte column, are 512 contiguous bytes.  Every thread
 * of a launch appends the same counts in the same order, so the bookkeeping
 * is uniform.  The completed K steps of the current super-step advance
 * through k0, k1, k2 (an indexed array would be placed in scratch memory). */
template <int S, bool GRP = false>
struct SuperStream {
    uint64_t acc;
    int32_t na, nq, np, pc3;
    int64_t u;
    uint32_t q0, q1, q2, q3;
    uint4 k0, k1, k2;

    __device__ __forceinline__ void init(int64_t u0) {
        acc = 0;
        na = nq = np = pc3 = 0;
        u = u0;
        q0 = q1 = q2 = q3 = 0;
        k0 = k1 = k2 = make_uint4(0u, 0u, 0u, 0u);
    }
    __device__ __forceinline__ void kstep(const uint4 &kr, uint4 *out, int64_t rows, int64_t r) {
        if (np == 0) {
            k0 = kr;
        } else if (np == 1) {
            k1 = kr;
        } else if (S == 4 && np == 2) {
            k2 = kr;
        } else {
            /* S = 3: kr is the third K step, bit 3 of every nibble stays 0 */
            const uint4 c = S == 4 ? k2 : kr, e = S == 4 ? kr : make_uint4(0u, 0u, 0u, 0u);
            if (r < rows) {
                /* GRP: uint4 g of the row is 32 uint4 after uint4 g - 1 */
                int64_t at, gs;
                if (GRP) {
                    const int64_t v = r / 3, pl = r - 3 * v;
                    at = (((u * (rows / 96) + (v >> 5)) * 3 + pl) * 4) * 32 + (v & 31);
                    gs = 32;
                } else {
                    at = (u * rows + r) * 4;
                    gs = 1;
                }
                uint4 *dst = out + at;
                dst[0] = super_group(k0.x, k1.x, c.x, e.x);
                dst[gs] = super_group(k0.y, k1.y, c.y, e.y);
                dst[2 * gs] = super_group(k0.z, k1.z, c.z, e.z);
                dst[3 * gs] = super_group(k0.w, k1.w, c.w, e.w);
            }
            pc3 += __popc(e.x) + __popc(e.y) + __popc(e.z) + __popc(e.w);
            u++;
        }
        np = np == S - 1 ? 0 : np + 1;
    }
    __device__ __forceinline__ void emit(uint32_t w, uint4 *out, int64_t rows, int64_t r) {
        q0 = q1;
        q1 = q2;
        q2 = q3;
        q3 = w;
        if (++nq == 4) {
            kstep(make_uint4(q0, q1, q2, q3), out, rows, r);
            nq = 0;
        }
    }
    __device__ __forceinline__ void append(uint32_t bits, int32_t k, uint4 *out, int64_t rows,
                                           int64_t r) {
        acc |= (uint64_t)bits << na;
        na += k;
        if (na >= 32) {
            emit((uint32_t)acc, out, rows, r);
            acc >>= 32;
            na -= 32;
        }
    }
    /* zero-pads the last K step and the last super-step of the stream */
    __device__ __forceinline__ void finish(uint4 *out, int64_t rows, int64_t r) {
        if (na > 0) {
            emit((uint32_t)acc, out, rows, r);
            acc = 0;
            na = 0;
        }
        while (nq != 0) emit(0u, out, rows, r);
        while (np != 0) kstep(make_uint4(0u, 0u, 0u, 0u), out, rows, r);
    }
};

/* One thread per plane row r = v*3 + p of [0, rows) (rows = n_var_pad * 3;
 * rows of padding variants v >= n_var are zero).  cw: [2][2*w64]
 * CompressWord, class 0 then class 1.  S K steps per super-step (4: single
 * scans, 3: the batch engine).  Class 0 fills super-steps [0, u1) (u1 =
 * f4_super or f4_super3 of its K steps), class 1 the super-steps from u1
 * on.  S = 4: pc3[c * rows + r] = set bits of row r in the fourth K steps
 * of class c's super-steps.  GRP: the grouped layout (SuperStream; rows a
 * multiple of 96).  The panel rows of a workgroup are staged
 * through LDS 8 words at a time (coalesced reads).                         */
constexpr int CLS_CHUNK = 8;
template <int S, bool GRP = false>
__global__ void __launch_bounds__(256) k_class_f4(const uint64_t *__restrict__ nat, int32_t n_var,
                                                   int32_t w64, const CompressWord *__restrict__ cw,
                                                   int64_t rows, int32_t u1,
                                                   uint4 *__restrict__ out,
                                                   int32_t *__restrict__ pc3) {
    __shared__ uint64_t srow[256][CLS_CHUNK + 1];
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int64_t r = r0 + threadIdx.x;
    const int64_t real_rows = (int64_t)n_var * 3;
    SuperStream<S, GRP> s0, s1;
    s0.init(0);
    s1.init(u1);
    const CompressWord *c0 = cw, *c1 = cw + 2 * w64;
    for (int32_t base = 0; base < w64; base += CLS_CHUNK) {
        __syncthreads();
        for (int32_t idx = threadIdx.x; idx < 256 * CLS_CHUNK; idx += 256) {
            const int32_t rl = idx / CLS_CHUNK, wd = idx % CLS_CHUNK;
            const int64_t rr = r0 + rl;
            uint64_t val = 0;
            if (rr < real_rows && base + wd < w64) val = nat[rr * w64 + base + wd];
            srow[rl][wd] = val;
        }
        __syncthreads();
        const int32_t nw = w64 - base < CLS_CHUNK ? w64 - base : CLS_CHUNK;
        for (int32_t wd = 0; wd < nw; wd++) {
            const uint64_t x = srow[threadIdx.x][wd];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const uint32_t xh = (uint32_t)(x >> (32 * h));
                const int32_t j = 2 * (base + wd) + h;
                const CompressWord a = c0[j], b = c1[j];
                if (a.cnt) s0.append(compress_word(xh, a), (int32_t)a.cnt, out, rows, r);
                if (b.cnt) s1.append(compress_word(xh, b), (int32_t)b.cnt, out, rows, r);
            }
        }
    }
    s0.finish(out, rows, r);
    s1.finish(out, rows, r);
    if (S == 4 && r < rows) {
        pc3[r] = s0.pc3;
        pc3[rows + r] = s1.pc3;
    }
}

/* missing[v] = n_samples - popc(valid plane of variant v); 64 threads per
 * variant.  Samples beyond n_samples are never valid in a panel. */
static __global__ void __launch_bounds__(256) k_qa_variants(const uint64_t *__restrict__ nat,
                                                            int64_t n_var, int32_t w64,
                                                            int32_t n_samples,
                                                            int32_t *__restrict__ missing) {
    const int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (v >= n_var) return;
    const uint64_t *valid = nat + (v * 3 + 2) * w64;
    int32_t c = 0;
    for (int32_t w = lane; w < w64; w += 64) c += __popcll(valid[w]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) missing[v] = n_samples - c;
}

/* valid[s] += number of variants in [v0, v0 + chunk) with a call for sample
 * s; grid (ceil(n_samples / 256), ceil(n_var / chunk)).  missing = n_var -
 * valid is taken on the host. */
static __global__ void __launch_bounds__(256) k_qa_samples(const uint64_t *__restrict__ nat,
                                                           int64_t n_var, int32_t w64,
                                                           int32_t n_samples, int32_t chunk,
                                                           int32_t *__restrict__ valid_count) {
    const int32_t s = blockIdx.x * 256 + threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.y * chunk;
    const int64_t v1 = v0 + chunk < n_var ? v0 + chunk : n_var;
    if (s >= n_samples) return;
    const int32_t w = s >> 6;
    const uint64_t bit = 1ull << (s & 63);
    int32_t c = 0;
    for (int64_t v = v0; v < v1; v++) c += (nat[(v * 3 + 2) * w64 + w] & bit) ? 1 : 0;
    if (c) atomicAdd(&valid_count[s], c);
}

/* dst variant i = src variant idx[i]: its three plane rows of w64 words. */
static __global__ void __launch_bounds__(256) k_gather_variants(const uint64_t *__restrict__ src,
                                                                const int64_t *__restrict__ idx,
                                                                int64_t n_keep, int32_t w64,
                                                                uint64_t *__restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per = (int64_t)3 * w64;
    if (t >= n_keep * per) return;
    const int64_t i = t / per, o = t - i * per;
    dst[i * per + o] = src[idx[i] * per + o];
}

} // namespace gepi

#endif
