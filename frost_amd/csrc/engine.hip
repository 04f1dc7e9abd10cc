// Streaming FDMT + matched-filter search on gfx950. See engine.hpp for the
// rules and the device layout.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <tuple>

namespace fa {

namespace {

constexpr int kBlock       = 256;
constexpr int kInitPer     = 8;    // outputs per thread per init block
constexpr uint64_t kTileMax = kAutoTileMax; // automatic time tile bounds
constexpr uint64_t kTileMin = kAutoTileMin;
constexpr uint32_t kMaxBox = 4096; // widest filter (frost's boxcar_max bound)
constexpr uint32_t kMaxWidths = 32; // width trials per plane row
constexpr uint64_t kCandCapMax = uint64_t{1} << 28; // candidate buffer growth bound
constexpr uint64_t kSpillCap   = uint64_t{1} << 22; // carried form: host spill records
constexpr uint64_t kRingTileMin = kCarriedTileMin; // carried form: automatic tile bounds
constexpr uint64_t kRingTileMax = kCarriedTileMax;

// ---- kernels ---------------------------------------------------------------

// Ring position of a sample p samples after ring position q, p in [-R, 2R - q).
__device__ __forceinline__ int wrap_pos(int p, int R) {
    p = p < 0 ? p + R : p;
    return p >= R ? p - R : p;
}

// Ring position of absolute sample t in a ring of R samples.
__device__ __forceinline__ int ring_of(uint64_t t, int R) {
    return static_cast<int>(t % static_cast<uint64_t>(static_cast<uint32_t>(R)));
}

// Level 0: per channel, the left-fold sums x(t) + x(t-1) + ... + x(t-D) for
// each of its row delays D (ascending), over [t0 - hc, t0 + tile). The
// channel's row table is staged in LDS; each thread carries kInitPer running
// sums through the rows in delay order. Carried state (ring != 0): hc is 0
// and row r's sample t goes to ring position t mod R_r (the row table holds
// R_r in place of h).
constexpr int kInitRows = 64; // level-0 rows per channel held in LDS

__global__ void __launch_bounds__(kBlock)
fdmt_init_kernel(const float* __restrict__ x, uint64_t xstride, int64_t xcol0,
                 const int32_t* __restrict__ chan, const int64_t* __restrict__ rows,
                 float* __restrict__ out, int tile_len, int nchans,
                 const uint64_t* __restrict__ gulp_abs_p, int64_t tile_t0, int ring) {
    extern __shared__ float sx[];
    __shared__ int64_t s_base[kInitRows];
    __shared__ int32_t s_d[kInitRows], s_h[kInitRows], s_q[kInitRows];
    const uint64_t t0abs = ring ? *gulp_abs_p + static_cast<uint64_t>(tile_t0) : 0;
    for (int c = blockIdx.y; c < nchans; c += gridDim.y) {
        const int rf = chan[4 * c], nr = chan[4 * c + 1], hc = chan[4 * c + 2],
                  dm = chan[4 * c + 3];
        const int total = tile_len + hc;
        // Blocks start at indices congruent to hc mod 64, so each wave's
        // stores cover whole 256-byte lines of the rows (sample 0 of a row is
        // 64-float aligned); block 0 starts up to 64 indices early. Rings are
        // multiples of 64 long, so sample t0 sits at t0 mod 64 in every ring.
        const int align = ring ? static_cast<int>((64 - (t0abs & 63)) & 63) : (hc & 63);
        const int i0    = blockIdx.x * (kBlock * kInitPer) + align - 64;
        if (i0 >= total) continue;
        // xc[i] is the sample at t0 - hc + i; indices from -dm are held.
        const float* xc = x + (uint64_t)c * xstride + (xcol0 - hc);
        const int span  = kBlock * kInitPer + dm;
        __syncthreads();
        for (int j = threadIdx.x; j < span; j += kBlock) {
            const int gi = i0 - dm + j;
            const float v = xc[min(max(gi, -dm), total - 1)];
            sx[j]        = gi >= -dm && gi < total ? v : 0.0f;
        }
        float acc[kInitPer];
        int k = 0;
        for (int rb = 0; rb < nr; rb += kInitRows) {
            if (rb > 0) __syncthreads(); // previous batch of the row table consumed
            const int nb = min(kInitRows, nr - rb);
            for (int j = threadIdx.x; j < nb; j += kBlock) {
                const int64_t* rw = rows + 3 * (int64_t)(rf + rb + j);
                s_d[j]            = (int32_t)rw[0];
                s_base[j]         = rw[1];
                s_h[j]            = (int32_t)rw[2];
                s_q[j]            = ring ? ring_of(t0abs, (int32_t)rw[2]) : 0;
            }
            __syncthreads();
            // Thread tid holds samples i0 + 4 tid + 4 kBlock g + (0..3),
            // g < kInitPer / 4: whole 16-byte groups, stored as one vector
            // (o + i0 is 64-float aligned).
            static_assert(kInitPer % 4 == 0, "init samples come in groups of four");
            constexpr int kGroups = kInitPer / 4;
            if (rb == 0) {
#pragma unroll
                for (int q = 0; q < kInitPer; ++q)
                    acc[q] = sx[dm + 4 * threadIdx.x + 4 * kBlock * (q / 4) + (q % 4)];
            }
            for (int jr = 0; jr < nb; ++jr) {
                const int D = s_d[jr];
                for (; k < D; ++k) {
#pragma unroll
                    for (int q = 0; q < kInitPer; ++q)
                        acc[q] = acc[q] +
                                 sx[dm + 4 * threadIdx.x + 4 * kBlock * (q / 4) + (q % 4) - (k + 1)];
                }
                if (ring) {
                    // Ring row: sample t0 + i at (q + i) mod R; a group of
                    // four starts on a multiple of 4 and never straddles the
                    // wrap (R is a multiple of 64).
                    const int R = s_h[jr], q = s_q[jr];
                    float* o    = out + s_base[jr];
#pragma unroll
                    for (int gq = 0; gq < kGroups; ++gq) {
                        const int ib = i0 + 4 * threadIdx.x + 4 * kBlock * gq;
                        if (ib >= 0 && ib + 3 < total) {
                            typedef float f4 __attribute__((ext_vector_type(4)));
                            const f4 v4 = {acc[4 * gq], acc[4 * gq + 1], acc[4 * gq + 2], acc[4 * gq + 3]};
                            __builtin_nontemporal_store(v4, reinterpret_cast<f4*>(o + wrap_pos(q + ib, R)));
                        } else {
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int i = ib + j;
                                if (i >= 0 && i < total) o[wrap_pos(q + i, R)] = acc[4 * gq + j];
                            }
                        }
                    }
                    continue;
                }
                const int skip = hc - s_h[jr];
                float* o       = out + s_base[jr] - skip;
#pragma unroll
                for (int gq = 0; gq < kGroups; ++gq) {
                    const int ib = i0 + 4 * threadIdx.x + 4 * kBlock * gq;
                    if (ib >= skip && ib + 3 < total) {
                        // Streamed like the tree launches' stores.
                        typedef float f4 __attribute__((ext_vector_type(4)));
                        const f4 v4 = {acc[4 * gq], acc[4 * gq + 1], acc[4 * gq + 2], acc[4 * gq + 3]};
                        __builtin_nontemporal_store(v4, reinterpret_cast<f4*>(o + ib));
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int i = ib + j;
                            if (i < total && i >= skip) o[i] = acc[4 * gq + j];
                        }
                    }
                }
            }
        }
    }
}

// Levels in+1 .. out in one pass: out[r][i] = tree sum of in[src_k + i] over
// the row's K = 2^depth descendants, i < tile + h_r.
// Tree-order sum of v[O .. O+K): ((v0 + v1) + (v2 + v3)) + ...
template <int K, int O = 0, int N>
__device__ __forceinline__ float tree_sum(const float (&v)[N]) {
    if constexpr (K == 1) {
        return v[O];
    } else {
        return tree_sum<K / 2, O>(v) + tree_sum<K / 2, O + K / 2>(v);
    }
}

// One block: G consecutive rows of the launch over one time chunk of
// kBlock * EPT samples. Blocks go round-robin to the 8 XCDs by linear id, so
// the id is decoded as (xcd = id % 8, k = id / 8) and each XCD sweeps all
// row groups of chunk xcd, then of chunk xcd + 8, ...: consecutive rows share
// most of their input rows, which then stay in that XCD's L2.
template <int K, int EPT, int G>
__global__ void __launch_bounds__(kBlock)
fdmt_merge_kernel(const float* __restrict__ in, float* __restrict__ out,
                  const int64_t* __restrict__ prog, int nrows, int ngroups, int tile_len) {
    const uint32_t id    = blockIdx.x;
    const uint32_t k     = id >> 3;
    const uint32_t group = k % (uint32_t)ngroups;
    const uint32_t chunk = (id & 7U) + 8U * (k / (uint32_t)ngroups);
    const int base       = (int)chunk * (kBlock * EPT);
    const int r_end      = min((int)group * G + G, nrows);
    for (int r = (int)group * G; r < r_end; ++r) {
        const int64_t* p = prog + (int64_t)r * (K + 2);
        const int len    = tile_len + (int)p[K + 1];
        if (base >= len) continue;
        const float* src[K];
#pragma unroll
        for (int j = 0; j < K; ++j) src[j] = in + p[1 + j];
        float* o = out + p[0];
        float v[EPT][K];
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int i = base + threadIdx.x + q * kBlock;
            if (i < len) {
#pragma unroll
                for (int j = 0; j < K; ++j) v[q][j] = src[j][i];
            }
        }
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int i = base + threadIdx.x + q * kBlock;
            if (i < len) o[i] = tree_sum<K>(v[q]);
        }
    }
}

// Staged gather. One block: one row block over a run of consecutive time
// chunks of kChunkG samples (tile-relative tau). Waves 0-3 (producers) load
// each distinct input row's window for chunk c + 1 into registers and stage
// it in one half of a double-buffered LDS area while waves 4-7 (consumers)
// sum each output row's K descendants for chunk c from the other half in
// tree order and store them coalesced. Producers issue only loads and
// consumers only stores, so neither waits on the other's memory traffic;
// one barrier per chunk (LDS-only fences) hands the halves over.
constexpr int kChunkG  = 256;
constexpr int kSpreadG = 64; // max shift spread of one input row within a block
constexpr int kGatherThreads = 512;

struct GatherDims {
    int nblocks;
    int slices; // row-block slices spread over the XCDs (8 / time groups, at least 1)
    int nbs;    // row blocks per slice
    int tile_len;
    int tau_origin;
    int half;   // floats per LDS half
    int chunks; // consecutive chunks per block
    int zero_slot; // offset in each half of a chunk of zeros (zero-channel leaves), or -1
    // Level-0 runs (X0): x points at tile sample 0 of channel 0; input
    // samples exist for tile samples [xlo, xhi].
    const float* x;
    int xlo, xhi;
    // Carried state (the kernel's RING form: rows are rings, GatherLeaf::ring
    // and row record [3]): tile sample 0 is absolute sample
    // *gulp_abs_p + tile_t0.
    const uint64_t* gulp_abs_p;
    int64_t tile_t0;
};

constexpr int kGatherRowsMax = 128; // output rows per row block

constexpr int kX0Grid = 384; // grid elements per channel window (6 per lane)
constexpr int kX0Win  = 448; // window samples (grid + delays, 7 per lane)
constexpr int kX0Dmax = 16;  // largest level-0 delay formed from the input
constexpr int kX0LeafMax = 64; // input rows of one level-0 run (LDS slots of >= 256 floats)

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Raw buffer access (gfx9 resource word 3): byte offsets at or past
// 4 * nfloats, negative ones included, read 0 and drop stores. The range
// check applies to the per-lane (VGPR) offset only, not to an instruction
// immediate, so every offset is passed whole in a VGPR (made opaque so the
// compiler cannot split a constant part into the immediate).
constexpr int kBufWord3 = 0x00020000;
constexpr int kBufNone  = INT32_MIN; // an offset outside every buffer (reads 0, no fetch)

__device__ __forceinline__ int whole_offset(int off) {
    asm volatile("" : "+v"(off));
    return off;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const float* base, int nfloats) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, nfloats * 4,
                                             kBufWord3);
}

__device__ __forceinline__ float buffer_load(__amdgpu_buffer_rsrc_t rs, int byte_off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, whole_offset(byte_off),
                                                                          0, 0));
}

// Stores of level rows are streamed (non-temporal cache policy, aux bit 1 on
// gfx950): a row is read back only by a later launch, so its lines are not
// kept in L2 at the expense of the input rows that neighbouring blocks
// share.
constexpr int kStoreStream = 2;

__device__ __forceinline__ void buffer_store(__amdgpu_buffer_rsrc_t rs, int byte_off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), rs, whole_offset(byte_off), 0,
                                          kStoreStream);
}

__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int K, int UMAX, bool X0, bool RING>
__global__ void __launch_bounds__(kGatherThreads)
fdmt_gather_kernel(const float* __restrict__ in, float* __restrict__ out,
                   const GatherBlock* __restrict__ blocks, const GatherLeaf* __restrict__ leaves,
                   const int32_t* __restrict__ rows, const GatherChan* __restrict__ chans,
                   GatherDims g) {
    extern __shared__ float lds[];
    __shared__ float xwin[X0 ? 4 : 1][X0 ? kX0Win : 1]; // per producer wave
    __shared__ int s_rq[kGatherRowsMax]; // carried: ring position of tile sample 0 per output row
    constexpr int M     = (kChunkG + kSpreadG) / 64;
    // Work item (row block, time group): the row blocks are cut into
    // g.slices contiguous slices of g.nbs blocks (one slice unless the tile
    // has fewer than 8 time groups); virtual group vg = (time group) *
    // slices + slice. Block id goes to XCD id % 8, which sweeps the blocks
    // of virtual groups xcd, xcd + 8, ... in row order.
    const uint32_t id   = blockIdx.x;
    const uint32_t k    = id >> 3;
    const uint32_t vg   = (id & 7U) + 8U * (k / (uint32_t)g.nbs);
    const uint32_t blk  = (vg % (uint32_t)g.slices) * (uint32_t)g.nbs + k % (uint32_t)g.nbs;
    const uint32_t grp  = vg / (uint32_t)g.slices;
    if (blk >= (uint32_t)g.nblocks) return; // uniform: past the last row block
    const int tau_first = g.tau_origin + (int)grp * (kChunkG * g.chunks);
    const GatherBlock B = blocks[blk];
    const int lo_c      = max(0, (-B.hmax - tau_first) / kChunkG); // first chunk in range
    const int hi_c      = min(g.chunks, (g.tile_len - tau_first + kChunkG - 1) / kChunkG);
    if (tau_first + kChunkG * g.chunks <= -B.hmax || lo_c >= hi_c) return; // uniform
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (g.zero_slot >= 0) {
        // Zero-channel leaves read this chunk of zeros in either half (no
        // producer writes there); the first hand-over barrier orders it.
        for (int i = threadIdx.x; i < kChunkG; i += kGatherThreads) {
            lds[g.zero_slot + i]          = 0.0f;
            lds[g.half + g.zero_slot + i] = 0.0f;
        }
    }
    // Level-0 runs: the run's channel groups and input rows, read by the
    // producers at every chunk, wait in LDS.
    __shared__ GatherChan s_chan[X0 ? 16 : 1];
    __shared__ GatherLeaf s_leaf[X0 ? kX0LeafMax : 1];
    if constexpr (X0) {
        for (int j = threadIdx.x; j < B.nchan; j += kGatherThreads) s_chan[j] = chans[B.chan_first + j];
        for (int j = threadIdx.x; j < B.nleaves; j += kGatherThreads) s_leaf[j] = leaves[B.leaf_first + j];
        lds_barrier();
    }
    if (X0 && wave < 4) {
        // Producer of level-0 rows: per channel group, a window of input
        // samples (two chunks of loads in flight, register sets A and B) is
        // written to this wave's LDS window; the left-fold sums
        // x(t) + x(t-1) + ... + x(t-D) are accumulated in that order on the
        // channel's grid and each input row is emitted into its slot when its
        // delay D is reached (rows sorted by delay).
        constexpr int MX = kX0Win / 64, MG = kX0Grid / 64;
        float va[UMAX][MX], vb[UMAX][MX];
        auto load = [&](float (&v)[UMAX][MX], int c) {
            const int tau0 = tau_first + c * kChunkG;
#pragma unroll
            for (int u = 0; u < UMAX; ++u) {
                const int gi        = min(wave + 4 * u, B.nchan - 1);
                const GatherChan ch = s_chan[gi];
                const int tw        = tau0 - ch.xs - ch.dmax; // tile sample of window element 0
                const float* xr     = g.x + ch.xoff;
                if (tw >= g.xlo && tw + kX0Win - 1 <= g.xhi) {
                    const float* src = xr + tw;
#pragma unroll
                    for (int m = 0; m < MX; ++m) v[u][m] = src[lane + 64 * m];
                } else {
#pragma unroll
                    for (int m = 0; m < MX; ++m)
                        v[u][m] = xr[min(max(tw + lane + 64 * m, g.xlo), g.xhi)];
                }
            }
        };
        auto stage = [&](const float (&v)[UMAX][MX], int c) {
            const int tau0 = tau_first + c * kChunkG;
            float* buf     = lds + (c & 1) * g.half;
            float* xw      = xwin[wave];
#pragma unroll
            for (int u = 0; u < UMAX; ++u) {
                const int gi = wave + 4 * u;
                if (gi < B.nchan) {
                    const GatherChan ch  = s_chan[gi];
                    const GatherLeaf* lv = s_leaf + (ch.leaf_first - B.leaf_first);
                    wave_sync();
                    // v[u] through uniform branches (u is not a constant
                    // here: the channel loop is not unrolled).
#pragma unroll
                    for (int q = 0; q < UMAX; ++q) {
                        if (q == u) {
#pragma unroll
                            for (int m = 0; m < MX; ++m) xw[lane + 64 * m] = v[q][m];
                        }
                    }
                    wave_sync();
                    // Lane l holds grid elements 6l .. 6l + 5 and the kX0Dmax
                    // samples before them: xl[q] = x(tile sample of grid 6l,
                    // minus kX0Dmax, plus q), window element 6l + dmax - kX0Dmax + q.
                    float xl[MG + kX0Dmax];
                    const int e0 = MG * lane + ch.dmax - kX0Dmax;
#pragma unroll
                    for (int q = 0; q < MG + kX0Dmax; ++q) xl[q] = xw[max(e0 + q, 0)];
                    float acc[MG];
#pragma unroll
                    for (int m = 0; m < MG; ++m) acc[m] = xl[m + kX0Dmax];
                    // The group's rows in delay order; the first four are
                    // held in registers (fetched before the sums), later
                    // ones (rare) are fetched when reached.
                    int fslot[4], flen[4], fsh[4], fhc[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const GatherLeaf& q = lv[min(k, ch.nleaf - 1)];
                        fslot[k]            = q.slot;
                        flen[k]             = q.len;
                        fsh[k]              = q.shmax;
                        fhc[k]              = q.hc;
                    }
                    int lk     = 0;
                    auto emit  = [&]() {
                        // lk is wave-uniform: select chains keep the fields in
                        // registers; rows past the fourth are fetched (rare).
                        auto pick = [&](const int (&f)[4]) {
                            return lk == 0 ? f[0] : lk == 1 ? f[1] : lk == 2 ? f[2] : f[3];
                        };
                        GatherLeaf lf{};
                        if (lk < 4) {
                            lf.slot  = pick(fslot);
                            lf.len   = pick(flen);
                            lf.shmax = pick(fsh);
                            lf.hc    = pick(fhc);
                        } else {
                            lf = lv[lk];
                        }
                        ++lk;
                        const int j0  = MG * lane - (ch.xs - lf.shmax); // slot index of grid 6l
                        const int tp0 = tau0 - ch.xs + MG * lane;
                        const int ts  = tau0 - lf.shmax; // tile sample of slot element 0
                        if (ts >= -lf.hc && ts + lf.len <= g.tile_len) {
#pragma unroll
                            for (int m = 0; m < MG; ++m) {
                                const int j = j0 + m;
                                if (j >= 0 && j < lf.len) buf[lf.slot + j] = acc[m];
                            }
                        } else {
#pragma unroll
                            for (int m = 0; m < MG; ++m) {
                                const int j = j0 + m, tp = tp0 + m;
                                if (j >= 0 && j < lf.len)
                                    buf[lf.slot + j] =
                                        (tp >= -lf.hc && tp < g.tile_len) ? acc[m] : 0.0f;
                            }
                        }
                    };
                    if (ch.dmask & 1U) emit();
                    // Fully unrolled (a uniform test per delay instead of a
                    // break), so every xl index is a constant register.
#pragma unroll
                    for (int d = 1; d <= kX0Dmax; ++d) {
                        if (d <= ch.dmax) {
#pragma unroll
                            for (int m = 0; m < MG; ++m) acc[m] = acc[m] + xl[m + kX0Dmax - d];
                            if ((ch.dmask >> d) & 1U) emit();
                        }
                    }
                }
            }
        };
        load(va, lo_c);
        if (lo_c + 1 < hi_c) load(vb, lo_c + 1);
        for (int c = lo_c; c < hi_c; c += 2) {
            stage(va, c);
            lds_barrier();
            if (c + 2 < hi_c) load(va, c + 2);
            if (c + 1 < hi_c) {
                stage(vb, c + 1);
                lds_barrier();
                if (c + 3 < hi_c) load(vb, c + 3);
            }
        }
    } else if (wave < 4) {
        // Producer: two chunks of loads in flight (register sets A and B).
        // Each input row is read through a buffer resource bounded to its
        // samples [t0 - hc, t0 + tile): loads outside it return 0, which is
        // the value staged for samples outside the row. The load count is
        // static, so waiting for set A leaves set B in flight.
        float va[UMAX][M], vb[UMAX][M];
        // Carried state: ring position of tile sample 0 in each input row.
        int rq[UMAX];
        if constexpr (RING) {
            const uint64_t t0abs = *g.gulp_abs_p + static_cast<uint64_t>(g.tile_t0);
#pragma unroll
            for (int u = 0; u < UMAX; ++u)
                rq[u] = ring_of(t0abs, leaves[B.leaf_first + min(wave + 4 * u, B.nleaves - 1)].ring);
        }
        auto load = [&](float (&v)[UMAX][M], int c) {
            const int tau0 = tau_first + c * kChunkG;
#pragma unroll
            for (int u = 0; u < UMAX; ++u) {
                const int sidx      = min(wave + 4 * u, B.nleaves - 1);
                const GatherLeaf lf = leaves[B.leaf_first + sidx];
                if constexpr (RING) {
                    // Ring row: tile sample t at (rq + t) mod R; samples at
                    // or past the tile end read 0 (they feed only dropped
                    // outputs); every read sample is at most the row's
                    // look-back before the tile.
                    const auto rs = row_rsrc(in + lf.src, lf.ring);
                    const int tb  = tau0 - lf.shmax + lane;
#pragma unroll
                    for (int m = 0; m < M; ++m) {
                        const int t   = tb + 64 * m;
                        const bool ok = t < g.tile_len && (m < M - 1 || lane + 64 * m < lf.len);
                        v[u][m]       = buffer_load(rs, ok ? wrap_pos(rq[u] + t, lf.ring) * 4 : kBufNone);
                    }
                    continue;
                }
                const auto rs       = row_rsrc(in + (lf.src + lf.shmax - lf.hc), lf.hc + g.tile_len);
                const int off0      = (tau0 - lf.shmax + lf.hc + lane) * 4;
#pragma unroll
                for (int m = 0; m < M - 1; ++m) v[u][m] = buffer_load(rs, off0 + 256 * m);
                // The last segment only holds the shift spread (len - 256
                // samples): other lanes read out of range, which fetches nothing.
                const bool tail    = lane + 64 * (M - 1) < lf.len;
                v[u][M - 1]        = buffer_load(rs, tail ? off0 + 256 * (M - 1) : kBufNone);
            }
        };
        auto stage = [&](const float (&v)[UMAX][M], int c) {
            float* buf = lds + (c & 1) * g.half;
#pragma unroll
            for (int u = 0; u < UMAX; ++u) {
                const int sidx = wave + 4 * u;
                if (sidx < B.nleaves) {
                    const GatherLeaf lf = leaves[B.leaf_first + sidx];
                    float* dst          = buf + lf.slot + lane;
#pragma unroll
                    for (int m = 0; m < M; ++m)
                        if (m < M - 1 || lane + 64 * m < lf.len) dst[64 * m] = v[u][m];
                }
            }
        };
        load(va, lo_c);
        if (lo_c + 1 < hi_c) load(vb, lo_c + 1);
        for (int c = lo_c; c < hi_c; c += 2) {
            stage(va, c);
            lds_barrier();
            if (c + 2 < hi_c) load(va, c + 2);
            if (c + 1 < hi_c) {
                stage(vb, c + 1);
                lds_barrier();
                if (c + 3 < hi_c) load(vb, c + 3);
            }
        }
    } else {
        // Consumer: wave 4 + q sums rows q, q + 4, ... of the run; lane l
        // covers samples l, l + 64, l + 128, l + 192 of the chunk, so one
        // LDS address per (row, descendant) serves four reads. Stores go
        // through a buffer resource bounded to the row's samples
        // [t0 - h, t0 + tile), so samples outside it are dropped without
        // branches. The next row's record is requested before this row's
        // LDS reads.
        // Row record: [out lo, out hi, h, pad, delta_0 .. delta_{K-1}].
        constexpr int RS = 4 + K;
        const int q      = wave - 4;
        const int r_end  = B.row_first + B.nrows;
        if constexpr (RING) {
            // Ring position of tile sample 0 in each of this wave's rows
            // (rows q, q + 4, ...: lane i takes row q + 4 i).
            const uint64_t t0abs = *g.gulp_abs_p + static_cast<uint64_t>(g.tile_t0);
            const int rr         = B.row_first + q + 4 * lane;
            if (lane < kGatherRowsMax / 4 && rr < r_end)
                s_rq[rr - B.row_first] = ring_of(t0abs, rows[(int64_t)rr * RS + 3]);
            wave_sync();
        }
        for (int c = lo_c; c < hi_c; ++c) {
            lds_barrier();
            const float* buf = lds + (c & 1) * g.half + lane;
            const int tau0   = tau_first + c * kChunkG + lane;
            int r            = B.row_first + q;
            if (r >= r_end) continue;
            int32_t rec[RS], nxt[RS];
#pragma unroll
            for (int j = 0; j < RS; ++j) rec[j] = rows[(int64_t)r * RS + j];
            for (; r < r_end; r += 4) {
                const bool more = r + 4 < r_end;
                if (more) {
#pragma unroll
                    for (int j = 0; j < RS; ++j) nxt[j] = rows[(int64_t)(r + 4) * RS + j];
                }
                float w[4][K];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const float* p = buf + rec[4 + j];
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) w[s4][j] = p[64 * s4];
                }
                const int64_t o = (int64_t)(uint32_t)rec[0] | ((int64_t)rec[1] << 32);
                const int h     = rec[2];
                if constexpr (RING) {
                    // Ring row: tile samples [0, tile) only (h is 0).
                    const int R   = rec[3];
                    const int q0  = s_rq[r - B.row_first];
                    const auto rs = row_rsrc(out + o, R);
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) {
                        const int t = tau0 + 64 * s4;
                        buffer_store(rs, t < g.tile_len ? wrap_pos(q0 + t, R) * 4 : kBufNone,
                                     tree_sum<K>(w[s4]));
                    }
                } else {
                    const auto rs  = row_rsrc(out + (o - h), h + g.tile_len);
                    const int off0 = (tau0 + h) * 4;
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) buffer_store(rs, off0 + 256 * s4, tree_sum<K>(w[s4]));
                }
                if (more) {
#pragma unroll
                    for (int j = 0; j < RS; ++j) rec[j] = nxt[j];
                }
            }
        }
    }
}

// Node form of the first launch when level 0 is formed from the input: one
// block holds a whole sub-band's rows (up to kNodeChans channels, every
// level-0 row they read and all output rows over them) over a run of short
// chunks, so each channel window is loaded and folded once per chunk for the
// whole sub-band. Waves 0-7 (producers) own two channels each: they keep the
// channel windows of chunks c + 1 and c + 2 in flight while folding chunk c
// into its level-0 rows, x(t) + x(t-1) + ... + x(t-D) in that order, four
// grid positions per lane in registers, and write each row's slot in one
// half of a double-buffered LDS area as aligned quads (zero outside the
// row's samples [t0 - hc, t0 + T)). Waves 8-15 (consumers) sum each output
// row's K rows from the other half in tree order, one row per quarter-wave,
// and store it. Producers only load and consumers only store, so neither
// waits on the other's memory traffic; one barrier per chunk hands the
// halves over.
constexpr int kNodeChunk   = 64;   // samples per chunk: one per lane
constexpr int kNodeChans   = 16;   // channels per block: two per producer wave
constexpr int kNodeGrid    = 256;  // grid positions per channel (chunk + shift spread): four per lane
constexpr int kNodeWin     = kNodeGrid + kX0Dmax; // window floats per channel
constexpr int kNodeRows    = 256;  // output rows per block
constexpr int kNodeThreads = 1024; // 8 producer and 8 consumer waves

template <int K, bool RING>
__global__ void __launch_bounds__(kNodeThreads)
fdmt_node_kernel(float* __restrict__ out, const GatherBlock* __restrict__ blocks,
                 const GatherLeaf* __restrict__ leaves, const int32_t* __restrict__ rows,
                 const GatherChan* __restrict__ chans, GatherDims g) {
    // [row slots, zero chunk: g.half] x 2 halves, [windows: kNodeChans x kNodeWin],
    // [row records: kNodeRows x (K + 4, rounded to quads)], [ring positions: kNodeRows]
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const uint32_t blk  = blockIdx.x % (uint32_t)g.nblocks;
    const uint32_t grp  = blockIdx.x / (uint32_t)g.nblocks;
    const int tau_first = g.tau_origin + (int)grp * (kNodeChunk * g.chunks);
    const GatherBlock B = blocks[blk];
    const int lo_c      = max(0, (-B.hmax - tau_first) / kNodeChunk);
    const int hi_c      = min(g.chunks, (g.tile_len - tau_first + kNodeChunk - 1) / kNodeChunk);
    if (tau_first + kNodeChunk * g.chunks <= -B.hmax || lo_c >= hi_c) return; // uniform
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    typedef float f4 __attribute__((ext_vector_type(4)));
    if (g.zero_slot >= 0 && threadIdx.x < kNodeChunk) {
        lds[g.zero_slot + threadIdx.x]          = 0.0f;
        lds[g.half + g.zero_slot + threadIdx.x] = 0.0f;
    }
    if (wave < 8) {
        // Producer. Window of channel i for chunk c: element e is tile
        // sample tau0 - xs - dmax + e, e < (grid length) + dmax, stored at
        // window index kX0Dmax - dmax + e, so that grid position p (tile
        // sample tau0 - xs + p) minus delay d sits at index p + kX0Dmax - d
        // whatever the channel's dmax. Lane l loads e = l + 64 m.
        constexpr int XM = (kNodeWin + 63) / 64;
        float* xwin      = lds + 2 * g.half;
        struct Chan {
            int64_t xoff;
            int xs, dmax, xmin;
            uint32_t dmask;
        } cs[2];
        // Row k of channel s, in lane k: vpk = (LDS float index of grid
        // position 0, quad-aligned, plus 256) | (qlo / 4) << 15 |
        // (qhi / 4) << 22 for its quads of grid positions [qlo, qhi); vhc =
        // its computed look-back. tqmin / tqmax: the range of the tile
        // sample of grid position 0 in which no quad of any row needs zeros.
        int vpk[2], vhc[2], tqmin[2], tqmax[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const GatherChan ch = chans[B.chan_first + min(wave + 8 * s, B.nchan - 1)];
            cs[s]               = Chan{ch.xoff, ch.xs, ch.dmax, ch.xmin, ch.dmask};
            const GatherLeaf lf = leaves[ch.leaf_first + min(lane, ch.nleaf - 1)];
            const int plo       = ch.xs - lf.shmax;
            const int qlo = plo & ~3, qhi = (plo + lf.len + 3) & ~3;
            vpk[s] = (lf.slot - plo + 256) | (qlo >> 2) << 15 | (qhi >> 2) << 22;
            vhc[s] = lf.hc;
            int lo = INT32_MIN, hi = INT32_MAX;
            for (int k = 0; k < ch.nleaf; ++k) {
                lo = max(lo, __builtin_amdgcn_readlane(-lf.hc - qlo, k));
                hi = min(hi, __builtin_amdgcn_readlane(g.tile_len - qhi, k));
            }
            tqmin[s] = lo;
            tqmax[s] = hi;
        }
        float xa[2][XM], xb[2][XM]; // windows of two chunks in flight
        // A fixed number of loads per chunk (channel records clamped, every
        // lane loading), so waiting for one chunk leaves the next in flight.
        auto xload = [&](float (&xv)[2][XM], int c) {
            const int tau0 = tau_first + c * kNodeChunk;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const Chan& ch  = cs[s];
                const int tw    = tau0 - ch.xs - ch.dmax;
                const float* xr = g.x + ch.xoff;
#pragma unroll
                for (int m = 0; m < XM; ++m) xv[s][m] = xr[min(max(tw + lane + 64 * m, g.xlo), g.xhi)];
            }
        };
        auto xstore = [&](const float (&xv)[2][XM]) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int i = wave + 8 * s;
                if (i < B.nchan) {
                    const Chan& ch = cs[s];
                    const int wl   = kNodeChunk + ch.xs - ch.xmin + ch.dmax;
                    float* xw      = xwin + i * kNodeWin + kX0Dmax - ch.dmax;
#pragma unroll
                    for (int m = 0; m < XM; ++m)
                        if (lane + 64 * m < wl) xw[lane + 64 * m] = xv[s][m];
                }
            }
        };
        auto fold = [&](int c) {
            const int tau0 = tau_first + c * kNodeChunk;
            float* half    = lds + (c & 1) * g.half;
#pragma unroll 1
            for (int s = 0; s < 2; ++s) {
                const int i = wave + 8 * s;
                if (i >= B.nchan) break;
                const Chan ch  = s ? cs[1] : cs[0];
                const int lpk  = s ? vpk[1] : vpk[0];
                const int lhc  = s ? vhc[1] : vhc[0];
                const int tmin = s ? tqmin[1] : tqmin[0];
                const int tmax = s ? tqmax[1] : tqmax[0];
                // Lane l: grid positions 4l .. 4l + 3 (past the grid: nothing).
                const int gl = kNodeChunk + ch.xs - ch.xmin;
                if (4 * lane >= gl) continue;
                float xr[kX0Dmax + 4]; // window indices 4l .. 4l + kX0Dmax + 3
                const f4* xq = reinterpret_cast<const f4*>(xwin + i * kNodeWin) + lane;
#pragma unroll
                for (int q = 0; q < (kX0Dmax + 4) / 4; ++q) {
                    const f4 v = xq[q];
#pragma unroll
                    for (int k = 0; k < 4; ++k) xr[4 * q + k] = v[k];
                }
                float acc[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[k] = xr[kX0Dmax + k];
                const int tq      = tau0 - ch.xs; // tile sample of grid position 0
                const bool inside = tq >= tmin && tq <= tmax; // no row needs zeros (uniform)
                int li            = 0;
                auto emit         = [&]() {
                    const uint32_t pk = (uint32_t)__builtin_amdgcn_readlane(lpk, li);
                    const int base    = (int)(pk & 0x7fffU) - 256;
                    const int qlo = (int)(pk >> 15 & 0x7fU) << 2, qhi = (int)(pk >> 22 & 0x7fU) << 2;
                    // Only the lanes whose quad lies in the row's slot write
                    // (masked lanes take no LDS bandwidth).
                    const bool mine   = (uint32_t)(4 * lane - qlo) < (uint32_t)(qhi - qlo);
                    f4 v;
                    if (inside) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) v[k] = acc[k];
                    } else {
                        // Only quads at the row's ends of samples need zeros.
                        const int hc = __builtin_amdgcn_readlane(lhc, li);
                        const int t0 = tq + 4 * lane;
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            v[k] = (t0 + k >= -hc && t0 + k < g.tile_len) ? acc[k] : 0.0f;
                    }
                    ++li;
                    if (mine) static_cast<f4*>(__builtin_assume_aligned(half + base, 16))[lane] = v;
                };
                if (ch.dmask & 1U) emit();
                // Every step taken (past dmax the sums are unused), so every
                // xr index is a constant register.
#pragma unroll
                for (int d = 1; d <= kX0Dmax; ++d) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) acc[k] = acc[k] + xr[kX0Dmax + k - d];
                    if ((ch.dmask >> d) & 1U) emit();
                }
            }
        };
        // Chunk c: its windows (loaded two chunks earlier) are stored to
        // this wave's LDS windows (read back by this wave only), chunk c + 2
        // is requested into the freed registers, then c is folded.
        auto step = [&](float (&xv)[2][XM], int c) {
            xstore(xv);
            wave_sync();
            if (c + 2 < hi_c) xload(xv, c + 2);
            fold(c);
            wave_sync(); // window reads done before the next store
            lds_barrier();
        };
        xload(xa, lo_c);
        if (lo_c + 1 < hi_c) xload(xb, lo_c + 1);
        for (int c = lo_c; c < hi_c; c += 2) {
            step(xa, c);
            if (c + 1 < hi_c) step(xb, c + 1);
        }
    } else {
        // Consumer: wave 8 + q sums rows 4 (q + 8 k) + (lane / 16): each
        // quarter-wave one row, lane l the sample pairs 2 (l % 16) + {0, 1}
        // and 32 later of the chunk (four LDS reads per address; each
        // quarter-wave stores 128 contiguous bytes per store). Row record in LDS (copied
        // at block start): [K slot offsets, out lo, out hi, h, ring
        // length]; carried rows also keep the ring position of tile
        // sample 0.
        constexpr int RW  = K;                // offset words
        constexpr int RS  = RW + 4;           // record words
        constexpr int RSL = (RS + 3) & ~3;    // LDS record stride (whole quads)
        int* srec         = reinterpret_cast<int*>(lds + 2 * g.half + kNodeChans * kNodeWin);
        int* sq0          = srec + kNodeRows * RSL;
        const int q       = wave - 8;
        for (int i = threadIdx.x - 512; i < B.nrows * RS; i += 512) {
            const int r = i / RS, f = i % RS;
            // Global record: [out lo, out hi, h, ring length, offsets].
            srec[r * RSL + f] = rows[(int64_t)(B.row_first + r) * RS + (f < RW ? f + 4 : f - RW)];
        }
        if constexpr (RING) {
            const uint64_t t0abs = *g.gulp_abs_p + static_cast<uint64_t>(g.tile_t0);
            for (int r = threadIdx.x - 512; r < B.nrows; r += 512)
                sq0[r] = ring_of(t0abs, rows[(int64_t)(B.row_first + r) * RS + 3]);
        }
        const int sub = lane & 15;
        for (int c = lo_c; c < hi_c; ++c) {
            lds_barrier();
            const float* half = lds + (c & 1) * g.half + 2 * sub;
            const int t       = tau_first + c * kNodeChunk + 2 * sub; // tile sample of this lane's first
            for (int k = 0; 4 * (q + 8 * k) < B.nrows; ++k) {
                const int rr = 4 * (q + 8 * k) + (lane >> 4);
                if (rr >= B.nrows) continue;
                int rec[RSL];
                {
                    const int4* rp = reinterpret_cast<const int4*>(srec + rr * RSL);
#pragma unroll
                    for (int v = 0; v < RSL / 4; ++v) {
                        const int4 x = rp[v];
                        rec[4 * v] = x.x, rec[4 * v + 1] = x.y, rec[4 * v + 2] = x.z, rec[4 * v + 3] = x.w;
                    }
                }
                float w[4][K];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const float* p = half + rec[j];
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) w[s4][j] = p[(s4 & 1) + 32 * (s4 >> 1)];
                }
                float v[4];
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) v[s4] = tree_sum<K>(w[s4]);
                const int64_t o = (int64_t)(uint32_t)rec[RW] | ((int64_t)rec[RW + 1] << 32);
                if constexpr (RING) {
                    // Ring row: tile samples [0, tile) only (h is 0).
                    const int R  = rec[RW + 3];
                    const int q0 = sq0[rr];
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) {
                        const int ts = t + (s4 & 1) + 32 * (s4 >> 1);
                        if (ts < g.tile_len) __builtin_nontemporal_store(v[s4], out + o + wrap_pos(q0 + ts, R));
                    }
                } else {
                    const int h = rec[RW + 2];
                    float* dst  = out + o + t;
                    // Both branches use every sum, so the LDS reads all stay
                    // ahead of the branch.
                    if (t >= -h && t + 33 < g.tile_len) {
                        typedef float f2v __attribute__((ext_vector_type(2)));
                        __builtin_nontemporal_store(f2v{v[0], v[1]}, reinterpret_cast<f2v*>(dst));
                        __builtin_nontemporal_store(f2v{v[2], v[3]}, reinterpret_cast<f2v*>(dst + 32));
                    } else {
#pragma unroll
                        for (int s4 = 0; s4 < 4; ++s4) {
                            const int d4 = (s4 & 1) + 32 * (s4 >> 1);
                            if (t + d4 >= -h && t + d4 < g.tile_len) __builtin_nontemporal_store(v[s4], dst + d4);
                        }
                    }
                }
            }
        }
    }
}

template <int K, bool RING>
void launch_node(const MergeLaunch& ml, float* out, const GatherDims& g, int tl, hipStream_t s) {
    const unsigned ngrp   = ceil_div(tl - g.tau_origin, kNodeChunk * g.chunks);
    const uint64_t blocks = static_cast<uint64_t>(ngrp) * ml.nblocks;
    if (blocks * kNodeThreads >= (uint64_t{1} << 32))
        throw std::invalid_argument("FDMT merge grid too large; reduce the time tile");
    const size_t lds = (2 * static_cast<size_t>(g.half) + kNodeChans * kNodeWin +
                        kNodeRows * (((K + 4 + 3) & ~3) + 1)) * sizeof(float);
    fdmt_node_kernel<K, RING><<<static_cast<unsigned>(blocks), kNodeThreads, lds, s>>>(
        out, ml.blocks.get(), ml.leaves.get(), ml.rows.get(), ml.chans.get(), g);
}

template <int K, int UMAX, bool X0, bool RING>
void launch_gather_u(const MergeLaunch& ml, const float* in, float* out, const GatherDims& g,
                     int tl, hipStream_t s) {
    const unsigned ngrp   = ceil_div(tl - g.tau_origin, kChunkG * g.chunks);
    GatherDims gd = g;
    gd.slices = static_cast<int>(std::max(1U, std::min<unsigned>(8U / std::max(ngrp, 1U), ml.nblocks)));
    gd.nbs    = static_cast<int>(ceil_div(ml.nblocks, static_cast<unsigned>(gd.slices)));
    const uint64_t nvg    = static_cast<uint64_t>(ngrp) * gd.slices;
    const uint64_t blocks = (nvg + 7) / 8 * 8 * gd.nbs;
    if (blocks * kGatherThreads >= (uint64_t{1} << 32))
        throw std::invalid_argument("FDMT merge grid too large; reduce the time tile");
    fdmt_gather_kernel<K, UMAX, X0, RING><<<static_cast<unsigned>(blocks), kGatherThreads,
                                      2 * g.half * sizeof(float), s>>>(
        in, out, ml.blocks.get(), ml.leaves.get(), ml.rows.get(), ml.chans.get(), gd);
}

int default_gather_chunks() {
    static const int n = [] {
        const char* v = std::getenv("FROST_AMD_GATHER_CHUNKS");
        const int c   = v ? std::atoi(v) : 16;
        return (c >= 1 && c <= 64) ? c : 16;
    }();
    return n;
}

// Chunks per node-form block (FROST_AMD_NODE_CHUNKS, default 64: 4096 samples).
int node_chunks() {
    static const int n = [] {
        const char* v = std::getenv("FROST_AMD_NODE_CHUNKS");
        const int c   = v ? std::atoi(v) : 64;
        return (c >= 1 && c <= 256) ? c : 64;
    }();
    return n;
}

template <int K, bool RING>
void launch_gather_r(const MergeLaunch& ml, const float* in, float* out, const GatherDims& g,
                     int tl, hipStream_t s) {
    if (ml.from_input) {
        switch (ml.max_leaves) {
        case 16: return launch_gather_u<K, 4, true, RING>(ml, in, out, g, tl, s);
        case 32: return launch_gather_u<K, 8, true, RING>(ml, in, out, g, tl, s);
        default: throw std::logic_error("staged FDMT gather: unsupported channels per block");
        }
    }
    switch (ml.max_leaves) {
    case 16: return launch_gather_u<K, 4, false, RING>(ml, in, out, g, tl, s);
    case 32: return launch_gather_u<K, 8, false, RING>(ml, in, out, g, tl, s);
    case 48: return launch_gather_u<K, 12, false, RING>(ml, in, out, g, tl, s);
    case 64: return launch_gather_u<K, 16, false, RING>(ml, in, out, g, tl, s);
    default: throw std::logic_error("staged FDMT gather: unsupported input rows per block");
    }
}

// x0: tile sample 0 of channel 0 of the input; [xlo, xhi] the tile samples
// held by the input buffer (used only by launches reading level 0 from x).
template <int K>
void launch_gather(const MergeLaunch& ml, const float* in, float* out, int tl, const float* x0,
                   int xlo, int xhi, int chunks, bool ring, const uint64_t* gulp_abs_p,
                   int64_t tile_t0, hipStream_t s) {
    static_assert(kChunkG == 256, "four consumer waves cover one chunk");
    GatherDims g{};
    g.gulp_abs_p = gulp_abs_p;
    g.tile_t0    = tile_t0;
    g.nblocks    = static_cast<int>(ml.nblocks);
    g.tile_len   = tl;
    g.tau_origin = -((ml.hmax + 63) & ~63); // chunks start on 64-sample boundaries
    g.half       = static_cast<int>(ml.lds_floats);
    g.zero_slot  = ml.zero_slot;
    g.chunks     = chunks;
    g.x          = x0;
    g.xlo        = xlo;
    g.xhi        = xhi;
    if (ml.node) {
        g.chunks = node_chunks();
        if (ring) return launch_node<K, true>(ml, out, g, tl, s);
        return launch_node<K, false>(ml, out, g, tl, s);
    }
    if (ring) return launch_gather_r<K, true>(ml, in, out, g, tl, s);
    return launch_gather_r<K, false>(ml, in, out, g, tl, s);
}

struct MergeShape {
    int ept = 4, rows = 4;
};

MergeShape merge_shape(int K) {
    MergeShape m;
    m.ept  = K == 2 ? 8 : 4;
    m.rows = 4;
    if (const char* v = std::getenv("FROST_AMD_MERGE_EPT")) m.ept = std::atoi(v);
    if (const char* v = std::getenv("FROST_AMD_MERGE_ROWS")) m.rows = std::atoi(v);
    return m;
}

template <int K, int EPT, int G>
void launch_merge_shape(const float* in, float* out, const int64_t* prog, unsigned nrows, int tl,
                        int hmax, hipStream_t s) {
    const unsigned nchunks = ceil_div(tl + hmax, kBlock * EPT);
    const unsigned ngroups = ceil_div(nrows, G);
    const uint64_t blocks  = (uint64_t)((nchunks + 7) / 8) * 8 * ngroups;
    if (blocks * kBlock >= (uint64_t{1} << 32))
        throw std::invalid_argument("FDMT merge grid too large; reduce the time tile");
    fdmt_merge_kernel<K, EPT, G><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(
        in, out, prog, static_cast<int>(nrows), static_cast<int>(ngroups), tl);
}

template <int K>
void launch_merge(const float* in, float* out, const int64_t* prog, unsigned nrows, int tl,
                  int hmax, hipStream_t s) {
    const MergeShape m = merge_shape(K);
#define FA_MERGE_CASE(E, G)                                                                  \
    if (m.ept == E && m.rows == G) return launch_merge_shape<K, E, G>(in, out, prog, nrows, tl, hmax, s);
    FA_MERGE_CASE(4, 1) FA_MERGE_CASE(4, 4) FA_MERGE_CASE(4, 8)
    FA_MERGE_CASE(8, 1) FA_MERGE_CASE(8, 4) FA_MERGE_CASE(8, 8)
    FA_MERGE_CASE(16, 1) FA_MERGE_CASE(16, 4)
#undef FA_MERGE_CASE
    throw std::invalid_argument("unsupported FDMT merge shape (elements per thread, rows)");
}

struct SearchArgs {
    const float* buf;
    const SearchRow* rows;
    int nrows;
    const float* inv_cur;
    const float* inv_or;
    const float* floor_cur; // smallest window sum whose S/N exceeds thr, per (row, width)
    const uint32_t* widths;
    int nw;
    int kmax;
    int wm1;
    int tile_len;
    int64_t tile_t0;   // gulp-local start of the tile
    const uint64_t* gulp_abs_p; // device: absolute sample of gulp-local 0
    float thr;
    int mode;
    int trap;
    OpenRun* open;
    CandRecord* cands;
    unsigned long long* counter;
    uint64_t cap;
    CandRecord* spill;   // records cap .. cap + spill_cap - 1 (pinned host memory), or null
    uint64_t spill_cap;
    float* keep;       // optional plane-row store [prow][keep_stride]
    uint64_t keep_stride;
};

// ---- parallel run tracking and candidate compaction -------------------------
//
// One wave, one chunk of up to 256 samples, lane l holding chunk samples
// 4l .. 4l + 3 (those at or past clen take no part). A run is a maximal
// stretch of above-threshold samples, reported at the first sample below
// the threshold after it, with its first maximum as the peak. The chunk's
// runs come from one segmented scan over the samples: a sample below the
// threshold, or one above it right after a sample below, starts a segment;
// a segment carries its largest S/N (the first of equal ones), the sample
// and width index of that maximum, and its first sample. The run open
// before the chunk is the element ahead of sample 0. Each lane closes at
// most two runs in its four samples; the wave counts them with ballots,
// takes one block of record slots with one atomic add and places each
// lane's records at its mbcnt offset. The run still open at the chunk's
// last sample is the carried state afterwards.

struct ScanElem {
    float v;    // largest S/N of the segment so far
    uint32_t p; // (arg + 1) | (first + 1) << 9 | width << 18 | head << 24;
                // arg / first = -1: the carried run's peak / start
};
constexpr uint32_t kScanArg   = 0x1FFu;
constexpr uint32_t kScanFirst = 0x1FFu << 9;
constexpr uint32_t kScanW     = 0x3Fu << 18;
constexpr uint32_t kScanHead  = 1u << 24;

__device__ __forceinline__ ScanElem scan_comb(ScanElem L, ScanElem R) {
    // R follows L: a head in R starts a new segment; otherwise R's maximum
    // replaces L's only when strictly larger (the first maximum is kept).
    ScanElem o = L;
    if (R.p & kScanHead) {
        o = R;
    } else if (R.v > L.v) {
        o.v = R.v;
        o.p = (L.p & (kScanFirst | kScanHead)) | (R.p & (kScanArg | kScanW));
    }
    return o;
}

__device__ __forceinline__ ScanElem scan_shfl_up(ScanElem e, int d) {
    return ScanElem{__shfl_up(e.v, d), static_cast<uint32_t>(__shfl_up(static_cast<int>(e.p), d))};
}

__device__ __forceinline__ int lanes_below(uint64_t m) {
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
}

// Where a wave's candidate records go: the device buffer, then (carried
// state) the pinned host spill; the counter counts every record.
struct CandSink {
    CandRecord* cands;
    CandRecord* spill;
    unsigned long long* counter;
    uint64_t cap, spill_cap;
};

__device__ __forceinline__ void put_record(const CandSink& k, unsigned long long idx, const CandRecord& r) {
    if (idx < k.cap) k.cands[idx] = r;
    else if (idx < k.cap + k.spill_cap) k.spill[idx - k.cap] = r;
}

// First record slot of this wave's n records (n uniform): one atomic add by
// lane 0, broadcast to the wave.
__device__ __forceinline__ unsigned long long wave_slots(const CandSink& k, uint32_t n, int lane) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(k.counter, static_cast<unsigned long long>(n));
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(base));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(base >> 32));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Runs of one width (or of the per-sample best width, across_widths) over
// one chunk: sn the S/N per sample, ab whether it is above the threshold
// (false for cold samples and those at or past clen), aw the width index
// reported with it. abs0 is the absolute sample of chunk sample 0. all:
// every above-threshold sample is a record of its own and no run stays open.
// Returns whether a run is open after the chunk (uniform).
__device__ __forceinline__ uint32_t run_chunk(const CandSink& a, OpenRun* st, uint32_t prow, const float (&sn)[4],
                              const bool (&ab)[4], const uint32_t (&aw)[4], bool all, int clen,
                              uint64_t abs0, int lane) {
    const uint64_t b0 = __ballot(ab[0]), b1 = __ballot(ab[1]), b2 = __ballot(ab[2]),
                   b3 = __ballot(ab[3]);
    if (all) {
        const int cnt     = (int)ab[0] + (int)ab[1] + (int)ab[2] + (int)ab[3];
        const uint64_t c1 = __ballot(cnt >= 1), c2 = __ballot(cnt >= 2), c3 = __ballot(cnt >= 3),
                       c4 = __ballot(cnt >= 4);
        const uint32_t tot = __popcll(c1) + __popcll(c2) + __popcll(c3) + __popcll(c4);
        const unsigned long long base = wave_slots(a, tot, lane);
        unsigned long long idx = base + lanes_below(c1) + lanes_below(c2) + lanes_below(c3) +
                                 lanes_below(c4);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (ab[j]) {
                const uint64_t t = abs0 + (uint64_t)(4 * lane + j);
                put_record(a, idx++, CandRecord{t, t, t, sn[j], prow, aw[j], 0});
            }
        }
        return 0;
    }
    const OpenRun o  = *st;
    const bool prev0 = lane == 0 ? (o.open != 0) : ((b3 >> (lane - 1)) & 1ULL) != 0;
    const bool prev[4] = {prev0, ab[0], ab[1], ab[2]};
    ScanElem e[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t idx = (uint32_t)(4 * lane + j) + 1U;
        if (4 * lane + j >= clen) e[j] = ScanElem{-INFINITY, 0U};                // takes no part
        else if (!ab[j]) e[j] = ScanElem{-INFINITY, kScanHead};                 // below: a break
        else if (!prev[j]) e[j] = ScanElem{sn[j], idx | (idx << 9) | (aw[j] << 18) | kScanHead};
        else e[j] = ScanElem{sn[j], idx | (aw[j] << 18)};
    }
    // Inclusive scan of the lane aggregates, then the prefix ahead of each lane.
    ScanElem A = scan_comb(scan_comb(e[0], e[1]), scan_comb(e[2], e[3]));
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const ScanElem B = scan_shfl_up(A, d);
        if (lane >= d) A = scan_comb(B, A);
    }
    ScanElem X = scan_shfl_up(A, 1);
    if (lane == 0) X = ScanElem{-INFINITY, 0U};
    const ScanElem C = o.open ? ScanElem{o.snr, ((uint32_t)o.w << 18) | kScanHead}
                              : ScanElem{-INFINITY, kScanHead};
    ScanElem cur = scan_comb(C, X);
    // Runs closing at this lane's samples: the state just before the close.
    ScanElem r[2] = {cur, cur};
    int rc[2]     = {0, 0};
    int cnt       = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int idx = 4 * lane + j;
        if (idx < clen && !ab[j] && prev[j]) {
            if (cnt == 0) {
                r[0]  = cur;
                rc[0] = idx;
            } else {
                r[1]  = cur;
                rc[1] = idx;
            }
            ++cnt;
        }
        cur = scan_comb(cur, e[j]);
    }
    auto record = [&](const ScanElem& s, int64_t last) {
        // last: chunk index of the run's last sample (-1: the previous chunk's).
        const int arg   = (int)(s.p & kScanArg) - 1;
        const int first = (int)((s.p & kScanFirst) >> 9) - 1;
        CandRecord rec;
        rec.begin = first < 0 ? o.start : abs0 + (uint64_t)first;
        rec.peak  = arg < 0 ? o.peak : abs0 + (uint64_t)arg;
        rec.end   = last < 0 ? o.end : abs0 + (uint64_t)last;
        rec.snr   = s.v;
        rec.prow  = prow;
        rec.w     = (s.p & kScanW) >> 18;
        rec.pad   = 0;
        return rec;
    };
    const uint64_t c1 = __ballot(cnt >= 1), c2 = __ballot(cnt >= 2);
    const uint32_t tot = __popcll(c1) + __popcll(c2);
    if (tot != 0) {
        const unsigned long long base = wave_slots(a, tot, lane);
        const unsigned long long idx  = base + lanes_below(c1) + lanes_below(c2);
        if (cnt >= 1) put_record(a, idx, record(r[0], rc[0] - 1));
        if (cnt >= 2) put_record(a, idx + 1, record(r[1], rc[1] - 1));
    }
    // The state after the chunk: lane 63's (samples past clen take no part).
    const ScanElem fin{__shfl(cur.v, 63), static_cast<uint32_t>(__shfl(static_cast<int>(cur.p), 63))};
    const int last       = clen - 1;
    const uint64_t blast = (last & 3) == 0 ? b0 : (last & 3) == 1 ? b1 : (last & 3) == 2 ? b2 : b3;
    const uint32_t open  = (uint32_t)((blast >> (last >> 2)) & 1ULL);
    wave_sync();
    if (lane == 0) {
        OpenRun n = o;
        n.open    = static_cast<uint16_t>(open);
        if (open) {
            const CandRecord rec = record(fin, last);
            n.start              = rec.begin;
            n.peak               = rec.peak;
            n.end                = rec.end;
            n.snr                = rec.snr;
            n.w                  = static_cast<uint16_t>(rec.w);
        }
        *st = n;
    }
    wave_sync();
    return open;
}

// run_chunk on S/N values staged in LDS, four consecutive ones per lane,
// with the width index per sample (across_widths) or one width for all.
__device__ __forceinline__ uint32_t run_chunk_staged(const CandSink& k, float thr, OpenRun* st,
                                                     uint32_t prow, const float* s_sn,
                                                     const uint8_t* s_aw, uint32_t wfix, bool all,
                                                     int clen, uint64_t abs0) {
    const int lane = threadIdx.x & 63;
    float sn[4];
    bool ab[4];
    uint32_t aw[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = 4 * lane + j;
        sn[j]       = s_sn[i];
        ab[j]       = i < clen && sn[j] > thr;
        aw[j]       = s_aw ? s_aw[i] : wfix;
    }
    return run_chunk(k, st, prow, sn, ab, aw, all, clen, abs0, lane);
}

// A searched plane row: sample t0 + t of the tile, t in [-(W_max - 1), tile).
// Linear rows are read at in + (W_max - 1) + t; carried rows at ring
// position (t0 + t) mod R.
template <bool RING> struct RowRead {
    const float* p; // linear: sample t0; ring: the ring start
    int q = 0, R = 0;
    __device__ __forceinline__ RowRead(const SearchArgs& a, const SearchRow& row, uint64_t gabs) {
        if constexpr (RING) {
            p = a.buf + row.in;
            R = row.ring;
            q = ring_of(gabs + static_cast<uint64_t>(a.tile_t0), R);
        } else {
            p = a.buf + row.in + a.wm1;
        }
    }
    __device__ __forceinline__ float at(int t) const {
        if constexpr (RING) return p[wrap_pos(q + t, R)];
        else return p[t];
    }
};

// ---- ring-ladder search (any widths up to 4096) ------------------------------
//
// One block of 256 threads per plane row, one sample per thread per chunk of
// 256. Ladder level b (D_0 = row, D_b(t) = D_{b-1}(t) + D_{b-1}(t - 2^{b-1}))
// lives in an LDS ring of R_b = 256 + lookback_b floats, where lookback_b is
// the furthest level-b value any window or the next level reads behind the
// sample (the host computes it from the widths). Sample t of the tile sits at
// ring position (t - u0) mod R_b; a chunk overwrites only values older than
// every read. Warm-up chunks from u0 <= -(W_max - 1) fill the rings with the
// row's look-back; values covering samples before -(W_max - 1) are never read
// by a searched window.
constexpr int kRingChunk  = 256;
constexpr int kRingLevels = 13; // widths up to 4096

struct RingSearchArgs {
    SearchArgs a;
    int off[kRingLevels];  // LDS float offset of level b's ring
    int size[kRingLevels]; // its length R_b
    int u0;                // first chunk start (tile-relative, multiple of 256)
    int lds_floats;        // sum of R_b
};

__device__ __forceinline__ int ring_pos(int rel, int R) {
    // rel in (-R, 2R)
    rel = rel >= R ? rel - R : rel;
    return rel < 0 ? rel + R : rel;
}

struct Ring {
    const float* lds;
    const int* off;
    const int* size;
    const int* base;
    // D_b at (sample i of the chunk) - d
    __device__ __forceinline__ float at(int b, int i, int d) const {
        return lds[off[b] + ring_pos(base[b] + i - d, size[b])];
    }
    // Window sum of width W ending d samples before sample i: its set bits,
    // highest first.
    __device__ __forceinline__ float wsum(uint32_t W, int i, int d) const {
        int k     = 31 - __clz(W);
        float acc = at(k, i, d);
        int off   = d + (1 << k);
        for (--k; k >= 0; --k) {
            if ((W >> k) & 1U) {
                acc = acc + at(k, i, off);
                off += 1 << k;
            }
        }
        return acc;
    }
    __device__ __forceinline__ float value(uint32_t W, int i, int trap) const {
        float s = wsum(W, i, 0);
        if (trap && W >= 3) {
            const uint32_t e = (W + 1) / 4;
            s = (s - 0.5f * wsum(e, i, 0)) - 0.5f * wsum(e, i, (int)(W - e));
        }
        return s;
    }
};

template <bool RING>
__global__ void __launch_bounds__(kRingChunk) search_ring_kernel(RingSearchArgs ra) {
    extern __shared__ float lds[];
    const SearchArgs& a = ra.a;
    __shared__ OpenRun st[32];
    __shared__ float s_inv[32], s_invor[32];
    __shared__ uint32_t s_w[32];
    __shared__ int64_t s_warm[32];
    __shared__ float s_best[kRingChunk];
    __shared__ uint8_t s_arg[kRingChunk];
    __shared__ int s_off[kRingLevels], s_size[kRingLevels], s_base[kRingLevels];
    __shared__ uint32_t flags, s_open; // widths with a sample above the threshold / a run open
    const int i         = threadIdx.x;
    const SearchRow row = a.rows[blockIdx.x];
    const uint32_t prow = row.prow;
    const int nw        = a.nw;
    const bool across   = a.mode == FA_PEAK_ACROSS_WIDTHS;
    const uint64_t gabs = *a.gulp_abs_p;
    if (i < nw) {
        st[i]     = a.open[(uint64_t)prow * nw + i];
        s_inv[i]  = a.inv_cur[(uint64_t)prow * nw + i];
        s_invor[i] = a.inv_or[(uint64_t)prow * nw + i];
        s_w[i]    = a.widths[i];
        s_warm[i] = (int64_t)row.delay + a.widths[i] - 1;
    }
    if (i < kRingLevels) {
        s_off[i]  = ra.off[i];
        s_size[i] = ra.size[i];
        s_base[i] = 0;
    }
    if (i == 0) {
        flags  = 0;
        s_open = 0;
    }
    __syncthreads();
    if (i < nw && st[i].open) atomicOr(&s_open, 1U << i);
    for (int j = i; j < ra.lds_floats; j += kRingChunk) lds[j] = 0.0f;
    __syncthreads();
    const Ring R{lds, s_off, s_size, s_base};
    const CandSink sink{a.cands, a.spill, a.counter, a.cap, a.spill_cap};
    const RowRead<RING> src(a, row, gabs);
    const int tl     = a.tile_len;
    for (int u = ra.u0; u < tl; u += kRingChunk) {
        const int t   = u + i;
        const float v = (t >= -a.wm1 && t < tl) ? src.at(t) : 0.0f;
        lds[s_off[0] + ring_pos(s_base[0] + i, s_size[0])] = v;
        if (a.keep && t >= 0 && t < tl)
            a.keep[(uint64_t)prow * a.keep_stride + a.tile_t0 + t] = v;
        for (int b = 1; b <= a.kmax; ++b) {
            __syncthreads();
            const float x = R.at(b - 1, i, 0) + R.at(b - 1, i, 1 << (b - 1));
            lds[s_off[b] + ring_pos(s_base[b] + i, s_size[b])] = x;
        }
        __syncthreads();
        if (u >= 0) {
            const int clen  = min(kRingChunk, tl - u);
            const int64_t tg = a.tile_t0 + t; // gulp-local sample
            const int64_t ab = (int64_t)gabs + tg;
            // S/N of width w at this thread's sample; false when the
            // window is cold (before delay + W - 1 in absolute samples);
            // windows reaching into the previous gulp use the OR-mask noise.
            auto snr_of = [&](int w, float& sn) {
                if (i >= clen || ab < s_warm[w]) return false;
                const float inv = tg < s_warm[w] ? s_invor[w] : s_inv[w];
                sn              = R.value(s_w[w], i, a.trap) * inv;
                return true;
            };
            if (across) {
                bool any   = false;
                float best = 0.0f;
                int bw     = 0;
                for (int w = 0; w < nw; ++w) {
                    float sn;
                    if (!snr_of(w, sn)) continue;
                    if (!any || sn > best) {
                        best = sn;
                        bw   = w;
                    }
                    any = true;
                }
                s_best[i] = any ? best : -INFINITY;
                s_arg[i]  = static_cast<uint8_t>(bw);
                if (any && best > a.thr) atomicOr(&flags, 1U);
            } else {
                uint32_t bits = 0;
                for (int w = 0; w < nw; ++w) {
                    float sn;
                    if (snr_of(w, sn) && sn > a.thr) bits |= 1U << w;
                }
                if (bits) atomicOr(&flags, bits);
            }
            __syncthreads();
            // Run tracking on wave 0 (a parallel scan over the chunk, four
            // samples per lane), one width at a time: those with a sample
            // above the threshold or a run open.
            const uint64_t abs0 = gabs + (uint64_t)(a.tile_t0 + u);
            uint32_t need       = flags | s_open;
            if (across) {
                if (need && i < 64) {
                    const uint32_t op = run_chunk_staged(sink, a.thr, &st[0], prow, s_best, s_arg, 0,
                                                         false, clen, abs0);
                    if (i == 0) s_open = op;
                }
            } else {
                while (need) {
                    const int w = __builtin_ctz(need);
                    need &= need - 1U;
                    float sn;
                    s_best[i] = snr_of(w, sn) ? sn : -INFINITY;
                    __syncthreads();
                    if (i < 64) {
                        const uint32_t op = run_chunk_staged(sink, a.thr, &st[w], prow, s_best, nullptr,
                                                             (uint32_t)w, a.mode == FA_PEAK_ALL, clen,
                                                             abs0);
                        if (i == 0) s_open = op ? (s_open | (1U << w)) : (s_open & ~(1U << w));
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
            if (i == 0) flags = 0;
        }
        // Next chunk: every ring advances by one chunk.
        if (i < kRingLevels) {
            const int nb = s_base[i] + kRingChunk;
            s_base[i]    = nb >= s_size[i] ? nb - s_size[i] : nb;
        }
        __syncthreads();
    }
    if (i < nw) a.open[(uint64_t)prow * nw + i] = st[i];
}

// ---- register-ladder search (power-of-two widths up to 256) -----------------
//
// One wave per plane row. A chunk is 256 samples: lane l holds samples
// 4l..4l+3 of the chunk in registers. Ladder level k is built from level k-1
// shifted by 2^(k-1) samples: in-lane for shifts below 4, otherwise a wave
// shuffle from lane l - shift/4, taking the previous chunk's registers for
// lanes that fall before the chunk. The sums are the same additions, in the
// same order, as the LDS ladder (D_k(t) = D_{k-1}(t) + D_{k-1}(t - 2^(k-1))).

constexpr int kRegChunk = 256;

template <int S>
__device__ __forceinline__ float back(const float (&cur)[4], const float (&prev)[4], int j,
                                      int lane) {
    // Value at sample (4*lane + j - S) of the chunk, S >= 1.
    const int p  = j - S;                      // relative to 4*lane
    const int dl = (p >= 0) ? 0 : -((-p + 3) / 4); // floor(p / 4)
    const int jj = p - 4 * dl;                  // 0..3
    if (dl == 0) return cur[jj];
    // Destination lane L reads source lane L - d (current chunk) or, for
    // L < d, lane L - d + 64 of the previous chunk: choose at the source,
    // then rotate once.
    const int d     = -dl;
    const float sel = lane < 64 - d ? cur[jj] : prev[jj];
    return __shfl(sel, (lane - d) & 63);
}

template <int K, int KMAX, bool TRAP> struct LadderStep;

struct RegSearchArgs {
    SearchArgs a;
    bool ring;        // carried rows (ring addressing)
    uint32_t present; // bit k set when width 2^k is searched
    int widx[16];     // width index of level k
};

template <int K>
__device__ __forceinline__ float ladder_back(const float (&D)[9][4], const float (&P)[9][4], int j,
                                             int lane) {
    // D_{K-1} at t - 2^(K-1)
    switch (K) {
    case 1: return back<1>(D[0], P[0], j, lane);
    case 2: return back<2>(D[1], P[1], j, lane);
    case 3: return back<4>(D[2], P[2], j, lane);
    case 4: return back<8>(D[3], P[3], j, lane);
    case 5: return back<16>(D[4], P[4], j, lane);
    case 6: return back<32>(D[5], P[5], j, lane);
    case 7: return back<64>(D[6], P[6], j, lane);
    default: return back<128>(D[7], P[7], j, lane);
    }
}

template <int K>
__device__ __forceinline__ float trap_back(const float (&D)[9][4], const float (&P)[9][4], int j,
                                           int lane) {
    // D_{K-2} at t - 3 * 2^(K-2)  (second trapezoid edge, e = 2^(K-2))
    switch (K) {
    case 2: return back<3>(D[0], P[0], j, lane);
    case 3: return back<6>(D[1], P[1], j, lane);
    case 4: return back<12>(D[2], P[2], j, lane);
    case 5: return back<24>(D[3], P[3], j, lane);
    case 6: return back<48>(D[4], P[4], j, lane);
    case 7: return back<96>(D[5], P[5], j, lane);
    default: return back<192>(D[6], P[6], j, lane);
    }
}



template <int KMAX, bool TRAP, bool RING, bool ACROSS>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(TRAP ? 1 : 6))) search_reg_kernel(RegSearchArgs ra) {
    const SearchArgs& a = ra.a;
    __shared__ OpenRun s_st[kBlock / 64][16];
    __shared__ float s_inv[kBlock / 64][2][KMAX + 1]; // [current, OR-mask][ladder level]
    // A chunk's S/N values per ladder level, staged for run tracking.
    __shared__ float s_sn[kBlock / 64][ACROSS ? 1 : KMAX + 1][kRegChunk];
    __shared__ uint8_t s_aw[kBlock / 64][kRegChunk];  // across_widths: the width index per sample
    const int lane = threadIdx.x % 64;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int rowi = blockIdx.x * (kBlock / 64) + wave;
    if (rowi >= a.nrows) return;
    const SearchRow row = a.rows[rowi];
    const uint32_t prow = row.prow;
    const int nw        = a.nw;
    const uint64_t gabs = *a.gulp_abs_p;
    // Quick-test floors per ladder level (+inf: width not searched); the
    // noise scales, needed only where a test passes, wait in LDS.
    float flo[KMAX + 1];
#pragma unroll
    for (int k = 0; k <= KMAX; ++k) {
        const int w = ra.widx[k];
        flo[k]      = w >= 0 ? a.floor_cur[(uint64_t)prow * nw + w] : INFINITY;
        if (lane == 0) {
            s_inv[wave][0][k] = w >= 0 ? a.inv_cur[(uint64_t)prow * nw + w] : 0.0f;
            s_inv[wave][1][k] = w >= 0 ? a.inv_or[(uint64_t)prow * nw + w] : 0.0f;
        }
    }
    uint32_t open_mask = 0; // bit w: width index w has an open run
    {
        uint32_t mine = 0;
        if (lane < nw) {
            s_st[wave][lane] = a.open[(uint64_t)prow * nw + lane];
            mine             = s_st[wave][lane].open ? (1U << lane) : 0U;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) mine |= __shfl_xor(mine, off);
        open_mask = __builtin_amdgcn_readfirstlane(mine);
    }
    wave_sync();
    const int64_t warm_max = (int64_t)row.delay + (1 << KMAX) - 1;
    const RowRead<RING> src(a, row, gabs); // sample t0 + t
    const int tl           = a.tile_len;
    const CandSink sink{a.cands, a.spill, a.counter, a.cap, a.spill_cap};
    // Ladder levels of the current chunk (L0) and of the previous one (L1).
    float L0[9][4], L1[9][4], nx[4];
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) L0[k][j] = L1[k][j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = -kRegChunk + 4 * lane + j;
        nx[j]       = (i >= -a.wm1 && i < tl) ? src.at(i) : 0.0f;
    }
    // One chunk starting at tile sample u: its levels D from the samples in
    // nx (the next chunk's are requested meanwhile) and the previous chunk's
    // levels P; then the threshold tests and run tracking.
    auto chunk = [&](float (&D)[9][4], const float (&P)[9][4], int u) {
        const int base = u + 4 * lane;
#pragma unroll
        for (int j = 0; j < 4; ++j) D[0][j] = nx[j];
        if (u + 2 * kRegChunk <= tl) {
#pragma unroll
            for (int j = 0; j < 4; ++j) nx[j] = src.at(base + kRegChunk + j);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = base + kRegChunk + j;
                nx[j]       = i < tl ? src.at(i) : 0.0f;
            }
        }
        if (a.keep && u >= 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (base + j < tl)
                    a.keep[(uint64_t)prow * a.keep_stride + a.tile_t0 + base + j] = D[0][j];
        }
        const int64_t tc0 = a.tile_t0 + u;
        const bool fast   = tc0 >= warm_max && u + kRegChunk <= tl;
#pragma unroll
        for (int k = 1; k <= KMAX; ++k) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float bj;
                switch (k) {
                case 1: bj = ladder_back<1>(D, P, j, lane); break;
                case 2: bj = ladder_back<2>(D, P, j, lane); break;
                case 3: bj = ladder_back<3>(D, P, j, lane); break;
                case 4: bj = ladder_back<4>(D, P, j, lane); break;
                case 5: bj = ladder_back<5>(D, P, j, lane); break;
                case 6: bj = ladder_back<6>(D, P, j, lane); break;
                case 7: bj = ladder_back<7>(D, P, j, lane); break;
                default: bj = ladder_back<8>(D, P, j, lane); break;
                }
                D[k][j] = D[k - 1][j] + bj;
            }
        }
        // Window sum of width 2^k ending at sample j (the trapezoid takes off
        // half of each edge of (2^k + 1) / 4 samples).
        auto wsum = [&](int k, int j) -> float {
            float sum = D[k][j];
            if (TRAP && k >= 2) {
                float e2;
                switch (k) {
                case 2: e2 = trap_back<2>(D, P, j, lane); break;
                case 3: e2 = trap_back<3>(D, P, j, lane); break;
                case 4: e2 = trap_back<4>(D, P, j, lane); break;
                case 5: e2 = trap_back<5>(D, P, j, lane); break;
                case 6: e2 = trap_back<6>(D, P, j, lane); break;
                case 7: e2 = trap_back<7>(D, P, j, lane); break;
                default: e2 = trap_back<8>(D, P, j, lane); break;
                }
                sum = (sum - 0.5f * D[k - 2][j]) - 0.5f * e2;
            }
            return sum;
        };
        // One test for the whole chunk first: with no run open and no window
        // sum at its width's floor (the smallest sum whose S/N exceeds the
        // threshold), nothing below can emit or change state.
        bool quiet = u < 0; // the warm-up chunk before the tile is never searched
        if (!quiet && fast && open_mask == 0) {
            bool hit = false;
#pragma unroll
            for (int k = 0; k <= KMAX; ++k) {
                if (!((ra.present >> k) & 1U)) continue;
#pragma unroll
                for (int j = 0; j < 4; ++j) hit = hit | (wsum(k, j) >= flo[k]);
            }
            quiet = __ballot(hit) == 0;
        }
        const int clen      = min(kRegChunk, tl - u);
        const uint64_t abs0 = gabs + (uint64_t)tc0;
        if constexpr (ACROSS) {
            // One run per row: per sample the largest S/N over the widths
            // searched at it (the lowest width index on ties).
            if (u >= 0 && (!quiet || open_mask != 0)) {
                float best[4];
                uint32_t arg[4];
                bool any[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    best[j] = 0.0f;
                    arg[j]  = 0;
                    any[j]  = false;
                }
#pragma unroll
                for (int k = 0; k <= KMAX; ++k) {
                    if (!((ra.present >> k) & 1U)) continue;
                    const int w        = ra.widx[k];
                    const int64_t warm = (int64_t)row.delay + (int64_t)(1 << k) - 1;
                    const float invc   = s_inv[wave][0][k];
                    const float invo   = s_inv[wave][1][k];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float sum = wsum(k, j);
                        const int64_t t = tc0 + 4 * lane + j;
                        if (base + j >= tl || (int64_t)gabs + t < warm) continue;
                        const float sn = sum * (t < warm ? invo : invc);
                        if (!any[j] || sn > best[j]) {
                            best[j] = sn;
                            arg[j]  = (uint32_t)w;
                        }
                        any[j] = true;
                    }
                }
                bool pred = false;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float sn = any[j] ? best[j] : -INFINITY;
                    pred           = pred | (sn > a.thr);
                    s_sn[wave][0][4 * lane + j] = sn;
                    s_aw[wave][4 * lane + j]    = static_cast<uint8_t>(arg[j]);
                }
                if (__ballot(pred) != 0 || open_mask != 0) {
                    wave_sync();
                    open_mask = run_chunk_staged(sink, a.thr, &s_st[wave][0], prow, s_sn[wave][0],
                                                 s_aw[wave], 0, false, clen, abs0)
                                    ? 1U : 0U;
                    wave_sync();
                }
            }
        } else if (!quiet) {
            uint32_t need = 0; // ladder levels whose S/N values wait in LDS
#pragma unroll
            for (int k = 0; k <= KMAX; ++k) {
                if (!((ra.present >> k) & 1U)) continue;
                const int w      = ra.widx[k];
                const uint32_t W = 1U << k;
                float sums[4], sv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) sums[j] = wsum(k, j);
                bool pred = false;
                if (fast) {
                    // The floor test first; the S/N values are formed only
                    // where it passes or a run is open.
                    bool pre = false;
#pragma unroll
                    for (int j = 0; j < 4; ++j) pre = pre | (sums[j] >= flo[k]);
                    if (__ballot(pre) == 0 && !((open_mask >> w) & 1U)) continue;
                    const float inv = s_inv[wave][0][k];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        sv[j] = sums[j] * inv;
                        pred  = pred | (sv[j] > a.thr);
                    }
                } else {
                    const float invc = s_inv[wave][0][k];
                    const float invo = s_inv[wave][1][k];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int64_t t = tc0 + 4 * lane + j;
                        float sn        = -INFINITY;
                        if (base + j < tl) {
                            if (t < (int64_t)row.delay + (int64_t)W - 1) {
                                // Cold until the absolute sample passes the warm-up.
                                if ((int64_t)gabs + t >= (int64_t)row.delay + (int64_t)W - 1)
                                    sn = sums[j] * invo;
                            } else {
                                sn = sums[j] * invc;
                            }
                        }
                        sv[j] = sn;
                        pred  = pred | (sn > a.thr);
                    }
                }
                if (__ballot(pred) != 0 || ((open_mask >> w) & 1U)) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) s_sn[wave][k][4 * lane + j] = sv[j];
                    need |= 1U << k;
                }
            }
            // Run tracking of the staged levels, one at a time (one copy of
            // the tracker in the code).
            if (need) wave_sync();
            while (need) {
                const int k = __builtin_ctz(need);
                need &= need - 1U;
                const int w       = __popc(ra.present & ((1U << k) - 1U)); // width index of 2^k
                const uint32_t op = run_chunk_staged(sink, a.thr, &s_st[wave][w], prow, s_sn[wave][k],
                                                     nullptr, (uint32_t)w, a.mode == FA_PEAK_ALL,
                                                     clen, abs0);
                open_mask = op ? (open_mask | (1U << w)) : (open_mask & ~(1U << w));
            }
            wave_sync();
        }
    };
    for (int u = -kRegChunk; u < tl; u += kRegChunk) {
        chunk(L0, L1, u);
#pragma unroll
        for (int k = 0; k <= KMAX; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) L1[k][j] = L0[k][j];
    }
    wave_sync();
    if (lane < nw) a.open[(uint64_t)prow * nw + lane] = s_st[wave][lane];
}

template <int KMAX, bool TRAP>
void launch_reg(const RegSearchArgs& ra, unsigned blocks, hipStream_t s) {
    const bool across = ra.a.mode == FA_PEAK_ACROSS_WIDTHS;
    if (ra.ring) {
        if (across) search_reg_kernel<KMAX, TRAP, true, true><<<blocks, kBlock, 0, s>>>(ra);
        else search_reg_kernel<KMAX, TRAP, true, false><<<blocks, kBlock, 0, s>>>(ra);
    } else {
        if (across) search_reg_kernel<KMAX, TRAP, false, true><<<blocks, kBlock, 0, s>>>(ra);
        else search_reg_kernel<KMAX, TRAP, false, false><<<blocks, kBlock, 0, s>>>(ra);
    }
}

using RegLauncher = void (*)(const RegSearchArgs&, unsigned, hipStream_t);

template <bool TRAP> RegLauncher reg_launcher(int kmax) {
    switch (kmax) {
    case 0: return launch_reg<0, TRAP>;
    case 1: return launch_reg<1, TRAP>;
    case 2: return launch_reg<2, TRAP>;
    case 3: return launch_reg<3, TRAP>;
    case 4: return launch_reg<4, TRAP>;
    case 5: return launch_reg<5, TRAP>;
    case 6: return launch_reg<6, TRAP>;
    case 7: return launch_reg<7, TRAP>;
    default: return launch_reg<8, TRAP>;
    }
}

__global__ void set_u64_kernel(unsigned long long* p, unsigned long long v0,
                               unsigned long long v1) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        p[0] = v0;
        p[1] = v1;
    }
}

// Candidate count 0 and frontier "the sample after this gulp", from the
// gulp's start in device memory.
__global__ void begin_gulp_kernel(unsigned long long* ctr, const uint64_t* params, uint64_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctr[0] = 0;
        ctr[1] = params[0] + n;
    }
}

__global__ void frontier_kernel(const OpenRun* open, uint64_t n, unsigned long long* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && open[i].open) atomicMin(out, (unsigned long long)open[i].start);
}

__global__ void zero_floats_kernel(float* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void zero_open_kernel(OpenRun* open, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) open[i] = OpenRun{0, 0, 0, 0.0f, 0, 0};
}

__global__ void copy_cols_kernel(const float* __restrict__ src, uint64_t sstride,
                                 float* __restrict__ dst, uint64_t dstride, uint64_t ncols,
                                 uint64_t nrows) {
    for (uint64_t r = blockIdx.y; r < nrows; r += gridDim.y)
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncols;
             i += (uint64_t)gridDim.x * blockDim.x)
            dst[r * dstride + i] = src[r * sstride + i];
}

// ---- host helpers ------------------------------------------------------------

// The smallest float s with fl(s * inv) > thr (float multiply, round to
// nearest), one step lower still, or -inf where no such bound exists. For
// a finite inv > 0, fl(s * inv) is non-decreasing in s, so a window sum
// below the returned value cannot give an S/N above thr: the search's quick
// test "sum >= floor" never skips a window that the exact test
// "sum * inv > thr" passes (the step of margin keeps it so even if the two
// products differed in their last bit).
float detect_floor(float inv, float thr) {
    if (!(inv > 0.0f) || !std::isfinite(inv) || !std::isfinite(thr)) return -INFINITY;
    auto pass = [&](float s) { return s * inv > thr; };
    float s = thr / inv;
    if (!std::isfinite(s)) return -INFINITY;
    int steps = 0;
    if (pass(s)) {
        for (;;) {
            const float d = std::nextafter(s, -INFINITY);
            if (!pass(d)) break;
            s = d;
            if (++steps > 64) return -INFINITY;
        }
    } else {
        do {
            s = std::nextafter(s, INFINITY);
            if (++steps > 64 || !std::isfinite(s)) return -INFINITY;
        } while (!pass(s));
    }
    return std::nextafter(s, -INFINITY);
}

double boxcar_window_variance(int64_t a, int64_t W) {
    const int64_t lo = std::min(a, W);
    const double l   = static_cast<double>(lo);
    return 2.0 * ((l - 1.0) * l * (2.0 * l - 1.0) / 6.0) +
           static_cast<double>(std::llabs(a - W) + 1) * l * l;
}

double trapezoid_window_variance(int64_t a, int64_t W) {
    if (W < 3) return boxcar_window_variance(a, W);
    const int64_t e = (W + 1) / 4;
    std::vector<double> pre(W + 1, 0.0);
    for (int64_t j = 0; j < W; ++j)
        pre[j + 1] = pre[j] + ((j < e || j >= W - e) ? 0.5 : 1.0);
    double total = 0.0;
    for (int64_t m = 0; m < W + a - 1; ++m) {
        const int64_t hi = std::min(W - 1, m);
        const int64_t lo = std::max<int64_t>(0, m - a + 1);
        const double s   = pre[hi + 1] - pre[lo];
        total += s * s;
    }
    return total;
}

// FROST_AMD_TILE overrides the time tile; 0 means choose automatically.
uint64_t env_tile() {
    if (const char* v = std::getenv("FROST_AMD_TILE")) {
        const long long t = std::atoll(v);
        if (t > 0) return static_cast<uint64_t>(t);
    }
    return 0;
}

uint64_t round64(uint64_t n) { return (n + 63) & ~uint64_t{63}; }

// Floats taken by a level row of look-back h over a tile: the row's sample 0
// sits on a 64-float boundary, so the look-back and the tile are each rounded.
uint64_t row_span(int32_t h, uint64_t tile) {
    return round64(static_cast<uint64_t>(h)) + round64(tile);
}

} // namespace

// ---- Engine ------------------------------------------------------------------

Engine::Engine(const FdmtPlan& plan, const fa_boxcar_config& cfg, const uint8_t* kill_mask,
               int device, uint64_t tile, uint64_t arena_bytes, LevelState state)
    : plan_(plan), cfg_(cfg), device_(device) {
    require(cfg.max_boxcar >= 1, "max_boxcar must be >= 1");
    require(cfg.max_boxcar <= kMaxBox, "max_boxcar must be in [1, 4096]");
    require(std::isfinite(cfg.detect_thresh) && cfg.detect_thresh > 0.0f,
            "detect_thresh must be finite and > 0");
    require(cfg.peak_mode == FA_PEAK_PER_WIDTH || cfg.peak_mode == FA_PEAK_ACROSS_WIDTHS ||
                cfg.peak_mode == FA_PEAK_ALL,
            "unknown peak_mode");
    require(cfg.max_candidates >= 1, "max_candidates must be >= 1");
    require(plan.cfg.block_nsamps >= 1, "block_nsamps (gulp capacity) must be >= 1");
    widths_ = box_width_trials(cfg.max_boxcar, cfg.wtsp);
    require(widths_.size() <= kMaxWidths,
            "too many width trials (at most 32): raise wtsp or lower max_boxcar");
    kmax_ = 0;
    while ((1U << (kmax_ + 1)) <= widths_.back()) ++kmax_;
    wm1_ = widths_.back() - 1;
    pow2_widths_ = true;
    for (const uint32_t w : widths_) pow2_widths_ = pow2_widths_ && (w & (w - 1)) == 0;
    // Register ladder: power-of-two widths up to 256 (shifts within one
    // 256-sample chunk); other widths use the LDS ring ladder.
    // FROST_AMD_SEARCH=ring forces the ring ladder.
    const char* sv = std::getenv("FROST_AMD_SEARCH");
    reg_allowed_ = pow2_widths_ && widths_.back() <= 256;
    reg_search_  = reg_allowed_ && !(sv && (std::string(sv) == "ring" || std::string(sv) == "lds"));
    {
        // Ring lengths: level b holds one chunk plus the furthest look-back
        // read from it (2^b by level b + 1; each set bit's offset inside a
        // window sum; the trapezoid's shifted edge window).
        std::vector<int32_t> look(kmax_ + 1, 0);
        for (uint32_t b = 0; b < kmax_; ++b) look[b] = std::max<int32_t>(look[b], 1 << b);
        auto add = [&](uint32_t W, int32_t d0) {
            for (uint32_t b = 0; b <= kmax_; ++b)
                if ((W >> b) & 1U) {
                    const int32_t off = static_cast<int32_t>(W - (W % (2U << b)));
                    look[b]           = std::max(look[b], d0 + off);
                }
        };
        for (const uint32_t W : widths_) {
            add(W, 0);
            if (cfg.use_trapezoid && W >= 3) {
                const uint32_t e = (W + 1) / 4;
                add(e, 0);
                add(e, static_cast<int32_t>(W - e));
            }
        }
        ring_off_.assign(kmax_ + 1, 0);
        ring_size_.assign(kmax_ + 1, 0);
        ring_floats_ = 0;
        for (uint32_t b = 0; b <= kmax_; ++b) {
            ring_off_[b]  = ring_floats_;
            ring_size_[b] = 256 + look[b];
            ring_floats_ += ring_size_[b];
        }
    }
    cap_ = plan.cfg.block_nsamps;
    const uint64_t tile_req = tile ? tile : env_tile();
    tile_ = std::min<uint64_t>(cap_, tile_req ? tile_req : kTileMax);
    const uint64_t nch = plan.cfg.nchans;
    ctor_mask_.assign(nch, 1);
    if (kill_mask) ctor_mask_.assign(kill_mask, kill_mask + nch);
    prev_mask_ = ctor_mask_;

    // Tuning: FROST_AMD_FUSE (levels per launch, 1-4), FROST_AMD_MERGE=direct
    // (gather from global memory instead of LDS staging), FROST_AMD_GATHER_LDS
    // (LDS floats per staged block), FROST_AMD_INIT=init (store level 0 with
    // the init kernel instead of forming it inside the first gather launch
    // with whole sub-bands per block, the default; fused-runs: formed in the
    // row runs of the other launches), FROST_AMD_NODE_LDS (LDS floats of a
    // sub-band block's rows).
    if (const char* fv = std::getenv("FROST_AMD_FUSE")) {
        const long f = std::atol(fv);
        if (f >= 1 && f <= 4) max_fuse_ = static_cast<uint32_t>(f);
    }
    if (const char* mv = std::getenv("FROST_AMD_MERGE")) staged_merge_ = std::string(mv) != "direct";
    if (const char* iv = std::getenv("FROST_AMD_INIT")) {
        x0_fused_ = std::string(iv) == "fused" || std::string(iv) == "fused-runs";
        x0_node_  = std::string(iv) != "fused-runs";
    }
    if (const char* nv = std::getenv("FROST_AMD_NODE_LDS")) {
        const long f = std::atol(nv);
        if (f >= 2048 && f <= 15000) node_lds_ = static_cast<uint32_t>(f); // two halves fit 160 KB
    }
    x0_fused_ = x0_fused_ && staged_merge_;
    if (!staged_merge_) max_fuse_ = std::min<uint32_t>(max_fuse_, 3);
    if (const char* gv = std::getenv("FROST_AMD_GATHER_LDS")) {
        const long f = std::atol(gv);
        if (f >= 2048 && f <= 16384) gather_lds_ = static_cast<uint32_t>(f);
    }

    DeviceGuard dg(device_);
    if (state == LevelState::automatic) {
        if (const char* sv2 = std::getenv("FROST_AMD_FDMT_STATE")) {
            const std::string v(sv2);
            if (v == "carried") state = LevelState::carried;
            if (v == "recompute") state = LevelState::recompute;
        }
    }
    shared_ = arena_bytes > 0;
    require(!(shared_ && state == LevelState::carried),
            "the carried FDMT level state keeps per-beam rings; it does not run in a shared arena");
    // Carried rings are addressed by the staged gather only.
    auto go_carried = [&](uint64_t t) {
        carried_      = true;
        staged_merge_ = true;
        tile_         = t;
        build_programs();
    };
    // The largest power-of-two carried tile whose input and rings fit
    // `budget` bytes (0: none).
    auto carried_pick = [&](uint64_t budget) -> uint64_t {
        for (uint64_t t = kRingTileMax; t >= kRingTileMin; t >>= 1) {
            const uint64_t tt = std::min<uint64_t>(t, cap_);
            if (carried_need(tt) <= budget) return tt;
        }
        return 0;
    };
    if (state == LevelState::carried) {
        uint64_t t = std::min<uint64_t>(cap_, tile_req ? tile_req : kRingTileMin);
        if (!tile_req && !sizing()) {
            size_t free_b = 0, total_b = 0;
            FA_HIP(hipMemGetInfo(&free_b, &total_b));
            if (const uint64_t pick = carried_pick(free_b / 10 * 9)) t = pick;
        }
        go_carried(t);
    } else {
        build_programs();
    }
    if (carried_) {
        // Tile chosen above.
    } else if (shared_) {
        // The largest power-of-two tile (up to kTileMax, the capacity and a
        // requested tile) whose input and level buffers, laid out as in the
        // arena, fit it. The stored levels are re-planned for the tile
        // picked, so its need is checked again after the programs are
        // rebuilt.
        uint64_t pick = 0;
        for (uint64_t t = kTileMax; t >= kTileMin && pick == 0; t >>= 1) {
            const uint64_t tt = std::min<uint64_t>(t, cap_);
            if (tile_req && tt > tile_req) continue;
            if (layout_need(tt) > arena_bytes) continue;
            if (tt != tile_) {
                tile_ = tt;
                build_programs();
            }
            if (layout_need(tile_) <= arena_bytes) pick = tile_;
        }
        if (pick == 0)
            throw OutOfMemory("the FDMT input and level buffers need " +
                              std::to_string(layout_need(std::min<uint64_t>(cap_, kTileMin))) +
                              " bytes at the smallest time tile, more than the " +
                              std::to_string(arena_bytes) + "-byte shared arena");
    } else if (!tile_req && !sizing()) {
        // Automatic tile: the largest power of two up to kTileMax (and the
        // gulp capacity) whose input and two level buffers fit in 75 % of
        // the device memory free now. Fewer, longer tiles recompute less
        // look-back and launch fewer kernels per gulp.
        size_t free_b = 0, total_b = 0;
        FA_HIP(hipMemGetInfo(&free_b, &total_b));
        const uint64_t budget = free_b / 4 * 3;
        uint64_t pick         = 0;
        for (uint64_t t = kTileMax; t >= kTileMin; t >>= 1) {
            const uint64_t tt = std::min<uint64_t>(t, cap_);
            if (buffers_bound(tt) <= budget) {
                pick = tt;
                break;
            }
        }
        // Nothing fits at the smallest tile: a large beam. Its look-back
        // would be recomputed many times over per tile; carry the levels in
        // rings instead, at the largest tile that fits 90 % of free memory.
        uint64_t ring_tile = 0;
        if (pick == 0 && state == LevelState::automatic) ring_tile = carried_pick(free_b / 10 * 9);
        if (ring_tile) {
            go_carried(ring_tile);
        } else {
            if (pick == 0) pick = std::min<uint64_t>(cap_, kTileMin);
            tile_ = pick;
            build_programs();
        }
    }

    // Extended input (zero history) and level buffers: owned (two ping-pong
    // buffers, or the rings of the carried form), or laid out in the shared
    // arena as [x | level buffer 0 | level buffer 1].
    xstride_ = round64(hist_in_ + cap_);
    const uint64_t nx = plan_.nchans_fdmt; // input rows: the channels and the zero channel
    uint64_t len0 = 0, len1 = 0;
    for (uint32_t l = 0; l < plan_.nlevels; ++l) {
        if (level_buf_[l] == 0) len0 = std::max(len0, level_len_[l]);
        if (level_buf_[l] == 1) len1 = std::max(len1, level_len_[l]);
    }
    // The last level's search still reads its buffer when the next tile's
    // first launch writes its output: when the two share a buffer (level 0
    // formed inside the first launch and an odd count of stored levels),
    // the two buffers swap roles every other tile so that they overlap.
    // Owned buffers only; both then hold the largest level.
    {
        uint32_t first = 0;
        while (first < plan_.nlevels && level_buf_[first] < 0) ++first;
        const uint32_t L = plan_.nlevels - 1;
        flip_ = !carried_ && !shared_ && first < L && level_buf_[first] == level_buf_[L];
        if (flip_) len0 = len1 = std::max(len0, len1);
    }
    len1 = std::max<uint64_t>(len1, 1);
    if (carried_) {
        x_.resize(nx * xstride_);
        if (!sizing()) FA_HIP(hipMemset(x_.get(), 0, x_.bytes()));
        rings_.resize(std::max<uint64_t>(rings_floats_, 1));
        if (!sizing()) FA_HIP(hipMemset(rings_.get(), 0, rings_.bytes())); // samples before the stream: 0
        x_ptr_      = x_.get();
        buf_ptr_[0] = buf_ptr_[1] = rings_.get();
        arena_need_ = x_.bytes() + rings_.bytes();
    } else if (shared_) {
        auto up256 = [](uint64_t b) { return (b + 255) & ~uint64_t{255}; };
        buf_off_[0] = up256(nx * xstride_ * sizeof(float));
        buf_off_[1] = up256(buf_off_[0] + len0 * sizeof(float));
        arena_need_ = buf_off_[1] + len1 * sizeof(float);
        hist_store_.resize(std::max<uint64_t>(nch * hist_in_, 1));
        if (!sizing()) FA_HIP(hipMemset(hist_store_.get(), 0, hist_store_.bytes()));
    } else {
        x_.resize(nx * xstride_);
        if (!sizing()) FA_HIP(hipMemset(x_.get(), 0, x_.bytes()));
        buf_[0].resize(len0);
        buf_[1].resize(len1);
        x_ptr_      = x_.get();
        buf_ptr_[0] = buf_[0].get();
        buf_ptr_[1] = buf_[1].get();
        arena_need_ = x_.bytes() + buf_[0].bytes() + buf_[1].bytes();
    }
    const uint64_t nw = widths_.size();
    inv_cur_.resize((uint64_t)nprows_ * nw);
    inv_or_.resize((uint64_t)nprows_ * nw);
    floor_cur_.resize((uint64_t)nprows_ * nw);
    widths_dev_.resize(nw);
    upload(widths_dev_.get(), widths_.data(), nw * 4);
    open_.resize((uint64_t)nprows_ * nw);
    open_saved_.resize((uint64_t)nprows_ * nw);
    if (!sizing()) FA_HIP(hipMemset(open_.get(), 0, open_.bytes()));
    cand_cap_ = std::max<uint64_t>(4 * cfg.max_candidates, uint64_t{1} << 20);
    cands_.resize(cand_cap_);
    if (carried_ && !sizing()) {
        // A carried gulp cannot be searched again (its rings have moved on):
        // records past the device buffer go to pinned host memory.
        spill_cap_ = kSpillCap;
        FA_HIP(hipHostMalloc(reinterpret_cast<void**>(&spill_), spill_cap_ * sizeof(CandRecord),
                             hipHostMallocMapped));
        FA_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&spill_dev_), spill_, 0));
    }
    counters_.resize(2);
    if (!sizing()) FA_HIP(hipMemset(counters_.get(), 0, counters_.bytes()));
    params_.resize(2);
    if (!sizing()) FA_HIP(hipMemset(params_.get(), 0, params_.bytes()));
    if (const char* gv = std::getenv("FROST_AMD_GRAPH")) graph_mode_ = std::string(gv) == "1";
    level_seconds.assign(plan_.nlevels, 0.0);
    level_launches.assign(plan_.nlevels, 0.0);
    if (sizing()) return;
    FA_HIP(hipStreamCreateWithFlags(&search_stream_, hipStreamNonBlocking));
    int lds_max = 0;
    FA_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device_));
    const uint64_t lds_need = static_cast<uint64_t>(ring_floats_) * sizeof(float) + 4096;
    if (!reg_search_ && lds_need > (uint64_t)lds_max)
        throw std::invalid_argument(
            "width trials need more LDS than the device offers for the ring-ladder search; "
            "raise wtsp or lower max_boxcar");
    FA_HIP(hipEventCreate(&ev_begin_));
    FA_HIP(hipEventCreate(&ev_end_));
}

uint64_t Engine::buffers_bound(uint64_t tile) const {
    // x plus two level buffers each as long as the longest level, with the
    // arena's alignment slack.
    const uint64_t nch    = plan_.nchans_fdmt; // input rows (with the zero channel)
    const uint64_t xbytes = nch * round64(hist_in_ + cap_) * sizeof(float);
    uint64_t lmax         = 0;
    for (uint32_t l = 0; l < plan_.nlevels; ++l) {
        uint64_t len = 0;
        for (const int32_t h : hist_[l]) len += row_span(h, tile);
        lmax = std::max(lmax, len);
    }
    return xbytes + 2 * lmax * sizeof(float) + 512;
}

uint64_t Engine::layout_need(uint64_t tile) const {
    // [x | level buffer 0 | level buffer 1], each buffer on a 256-byte
    // boundary, with the levels stored in each buffer as planned now.
    auto up256        = [](uint64_t b) { return (b + 255) & ~uint64_t{255}; };
    const uint64_t nch = plan_.nchans_fdmt;
    uint64_t len[2]    = {0, 1};
    for (uint32_t l = 0; l < plan_.nlevels; ++l) {
        if (level_buf_[l] < 0) continue;
        uint64_t n = 0;
        for (const int32_t h : hist_[l]) n += row_span(h, tile);
        len[level_buf_[l]] = std::max(len[level_buf_[l]], n);
    }
    return up256(up256(nch * round64(hist_in_ + cap_) * sizeof(float)) + len[0] * sizeof(float)) +
           len[1] * sizeof(float);
}

void Engine::bind_arena(void* arena) {
    require(shared_, "bind_arena: the engine owns its buffers");
    require(arena != nullptr, "bind_arena: null arena");
    require(!pending_collect_, "bind_arena: the last gulp has not been collected");
    auto* base  = static_cast<unsigned char*>(arena);
    x_ptr_      = reinterpret_cast<float*>(base);
    buf_ptr_[0] = reinterpret_cast<float*>(base + buf_off_[0]);
    buf_ptr_[1] = reinterpret_cast<float*>(base + buf_off_[1]);
}

Engine::~Engine() {
    drop_graphs();
    if (capture_stream_) (void)hipStreamDestroy(capture_stream_);
    if (search_stream_) {
        (void)hipStreamSynchronize(search_stream_);
        (void)hipStreamDestroy(search_stream_);
    }
    for (auto e : sync_events_) (void)hipEventDestroy(e);
    for (auto e : events_) (void)hipEventDestroy(e);
    if (ev_begin_) (void)hipEventDestroy(ev_begin_);
    if (ev_end_) (void)hipEventDestroy(ev_end_);
    if (spill_) (void)hipHostFree(spill_);
}

void Engine::plane_setup() {
    plane_first_.clear();
    nprows_ = 0;
    for (const auto& p : plan_.planes) {
        plane_first_.push_back(nprows_);
        nprows_ += static_cast<uint32_t>(p.delay.size());
    }
    dmax0_ = 0;
    for (const int32_t d : plan_.levels[0].delay) dmax0_ = std::max(dmax0_, d);
    if (dmax0_ > kX0Dmax) x0_fused_ = false; // level-0 delays too long for the fused form
    for (const auto& p : plan_.planes)
        if (p.level == 0) x0_fused_ = false; // a searched level 0 must be stored
}

// Level-0 program: per channel [first row, rows, computed look-back, largest
// delay]; per row [D, offset, h] (carried rings: [D, ring start, R]). The
// input history H covers the earliest sample a level-0 value reads: its
// computed look-back (when level 0 is formed inside the first launch, the
// look-back that launch reads) plus its delay.
void Engine::build_init() {
    const auto& l0     = plan_.levels[0];
    const uint64_t nch = plan_.nchans_fdmt;
    const bool ring0   = carried_ && !x0_fused_;
    std::vector<int32_t> chan(4 * nch, 0);
    std::vector<int64_t> rows(3 * l0.size());
    for (std::size_t r = 0; r < l0.size(); ++r) {
        const int32_t c = l0.chan_or_lo[r];
        const int32_t h = ring0 ? 0 : hist_[0][r];
        if (r == 0 || l0.chan_or_lo[r - 1] != c) chan[4 * c] = static_cast<int32_t>(r);
        chan[4 * c + 1] += 1;
        chan[4 * c + 2] = std::max(chan[4 * c + 2], h);
        chan[4 * c + 3] = std::max(chan[4 * c + 3], l0.delay[r]);
        rows[3 * r]     = l0.delay[r];
        rows[3 * r + 1] = base_[0][r];
        rows[3 * r + 2] = ring0 ? ring_len_[0][r] : h;
    }
    hist_in_ = 0;
    for (uint64_t c = 0; c < nch; ++c)
        hist_in_ = std::max<uint64_t>(hist_in_, static_cast<uint64_t>(chan[4 * c + 2]) +
                                                    static_cast<uint64_t>(chan[4 * c + 3]));
    xstride_ = round64(hist_in_ + cap_); // extended input row: H history + capacity
    init_chan_.resize(chan.size());
    upload(init_chan_.get(), chan.data(), chan.size() * 4);
    init_rows_.resize(rows.size());
    upload(init_rows_.get(), rows.data(), rows.size() * 8);
}

void Engine::build_searches() {
    search_rows_.clear();
    search_rows_.resize(plan_.nlevels);
    search_count_.assign(plan_.nlevels, 0);
    for (uint32_t l = 0; l < plan_.nlevels; ++l) {
        std::vector<SearchRow> sr;
        for (std::size_t p = 0; p < plan_.planes.size(); ++p) {
            const auto& pd = plan_.planes[p];
            if (pd.level != l) continue;
            for (std::size_t i = 0; i < pd.delay.size(); ++i) {
                const int32_t r = pd.level_row[i];
                SearchRow s{};
                if (carried_) {
                    s.in   = base_[l][r];
                    s.ring = ring_len_[l][r];
                } else {
                    s.in   = base_[l][r] + (hist_[l][r] - static_cast<int32_t>(wm1_));
                    s.ring = 0;
                }
                s.delay = pd.delay[i];
                s.prow  = plane_first_[p] + static_cast<uint32_t>(i);
                sr.push_back(s);
            }
        }
        search_count_[l] = static_cast<uint32_t>(sr.size());
        if (sr.empty()) continue;
        search_rows_[l].resize(sr.size());
        upload(search_rows_[l].get(), sr.data(), sr.size() * sizeof(SearchRow));
    }
}

void Engine::build_programs() {
    if (carried_) {
        build_carried();
        return;
    }
    const uint32_t L = plan_.nlevels - 1;
    plane_setup();
    hist_.assign(plan_.nlevels, {});
    for (uint32_t l = 0; l <= L; ++l) hist_[l].assign(plan_.levels[l].size(), -1);
    // Look-back per row: planes need W_max - 1, upper halves their shift.
    for (int l = static_cast<int>(L); l >= 0; --l) {
        for (const auto& p : plan_.planes)
            if (p.level == static_cast<uint32_t>(l))
                for (const int32_t r : p.level_row)
                    hist_[l][r] = std::max<int32_t>(hist_[l][r], static_cast<int32_t>(wm1_));
        if (l == 0) break;
        const auto& lv = plan_.levels[l];
        for (std::size_t r = 0; r < lv.size(); ++r) {
            const int32_t h = hist_[l][r];
            if (h < 0) throw std::logic_error("FDMT plan has an unused row");
            auto& lo = hist_[l - 1][lv.chan_or_lo[r]];
            auto& hi = hist_[l - 1][lv.hi[r]];
            lo       = std::max(lo, h);
            hi       = std::max(hi, h + lv.shift[r]);
        }
    }
    // Buffer layout per level.
    base_.assign(plan_.nlevels, {});
    ring_len_.assign(plan_.nlevels, {});
    level_len_.assign(plan_.nlevels, 0);
    for (uint32_t l = 0; l <= L; ++l) {
        uint64_t off = 0;
        // Rows start h samples before their sample 0, which is placed on a
        // 64-float boundary (whole 256-byte lines for every wave's stores).
        for (const int32_t h : hist_[l]) {
            base_[l].push_back(static_cast<int64_t>(off + round64(static_cast<uint64_t>(h)) - h));
            off += row_span(h, tile_);
        }
        level_len_[l] = off;
    }
    build_init();
    // Stored levels: level 0, every level holding a plane, and enough others
    // that no launch spans more than max_fuse_ levels; the fewest stored
    // floats win (each stored level is written and read once per tile).
    std::vector<bool> must(plan_.nlevels, false);
    must[0] = must[L] = true;
    for (const auto& p : plan_.planes) must[p.level] = true;
    const uint32_t F = std::max<uint32_t>(1, std::min<uint32_t>(max_fuse_, 4));
    constexpr uint64_t kInf = ~uint64_t{0};
    std::vector<uint64_t> best(plan_.nlevels, kInf);
    std::vector<int32_t> prev(plan_.nlevels, -1);
    best[0] = x0_fused_ ? 0 : level_len_[0]; // fused: level 0 is formed from the input
    for (uint32_t l = 1; l <= L; ++l) {
        for (uint32_t j = l; j-- > 0 && l - j <= F;) {
            if (best[j] != kInf && best[j] + level_len_[l] < best[l]) {
                best[l] = best[j] + level_len_[l];
                prev[l] = static_cast<int32_t>(j);
            }
            if (must[j]) break; // a plane level cannot be skipped
        }
    }
    std::vector<uint32_t> stored;
    for (int32_t l = static_cast<int32_t>(L); l > 0; l = prev[l]) {
        if (l < 0 || prev[l] < 0) throw std::logic_error("FDMT level fusion has no solution");
        stored.push_back(static_cast<uint32_t>(l));
    }
    stored.push_back(0);
    std::reverse(stored.begin(), stored.end());
    level_source_.assign(plan_.nlevels, -2);
    level_buf_.assign(plan_.nlevels, -1);
    level_source_[0] = x0_fused_ ? -2 : -1;
    for (std::size_t i = 0, k = 0; i < stored.size(); ++i) {
        if (stored[i] == 0 && x0_fused_) continue;
        level_buf_[stored[i]] = static_cast<int>(k++ & 1);
    }
    launches_.clear();
    for (std::size_t i = 1; i < stored.size(); ++i)
        add_launch(stored[i - 1], stored[i]);
    build_searches();
}

// Descendants of (level l, row r) at level a in tree order, with the summed
// shift of the upper branches taken on the way down.
static void tree_leaves(const FdmtPlan& plan, uint32_t a, uint32_t l, int32_t r, int64_t sh,
                        std::vector<std::pair<int32_t, int64_t>>& out) {
    if (l == a) {
        out.emplace_back(r, sh);
        return;
    }
    const auto& v = plan.levels[l];
    tree_leaves(plan, a, l - 1, v.chan_or_lo[r], sh, out);
    tree_leaves(plan, a, l - 1, v.hi[r], sh + v.shift[r], out);
}

// The launch computing level b from stored level a: its direct program
// [out, src_0 .. src_{K-1}, h] (linear rows only) and its staged program.
void Engine::add_launch(uint32_t a, uint32_t b) {
    level_source_[b] = static_cast<int32_t>(a);
    MergeLaunch ml;
    ml.in_level   = a;
    ml.out_level  = b;
    ml.radix      = 1 << (b - a);
    ml.in_buf     = level_buf_[a];
    ml.out_buf    = level_buf_[b];
    ml.from_input = a == 0 && x0_fused_;
    const auto& lv = plan_.levels[b];
    ml.nrows       = static_cast<uint32_t>(lv.size());
    const int K    = ml.radix;
    std::vector<GatherRow> orow(ml.nrows);
    std::vector<std::pair<int32_t, int32_t>> rl(static_cast<std::size_t>(ml.nrows) * K);
    std::vector<int64_t> prog;
    if (!carried_) prog.resize(static_cast<std::size_t>(ml.nrows) * (K + 2));
    std::vector<std::pair<int32_t, int64_t>> leaves;
    for (uint32_t r = 0; r < ml.nrows; ++r) {
        // Carried rows compute only the tile; linear rows their look-back too.
        const int32_t h = carried_ ? 0 : hist_[b][r];
        ml.hmax         = std::max(ml.hmax, h);
        leaves.clear();
        tree_leaves(plan_, a, b, static_cast<int32_t>(r), 0, leaves);
        orow[r] = GatherRow{carried_ ? base_[b][r] : base_[b][r] + h, h,
                            carried_ ? ring_len_[b][r] : 0};
        for (int k = 0; k < K; ++k) {
            const auto [c, sh] = leaves[k];
            if (hist_[a][c] < h + sh) throw std::logic_error("FDMT look-back too short");
            rl[static_cast<std::size_t>(r) * K + k] = {c, static_cast<int32_t>(sh)};
            if (!carried_)
                prog[static_cast<std::size_t>(r) * (K + 2) + 1 + k] =
                    base_[a][c] + (hist_[a][c] - h - sh);
        }
        if (!carried_) {
            prog[static_cast<std::size_t>(r) * (K + 2)]         = base_[b][r];
            prog[static_cast<std::size_t>(r) * (K + 2) + K + 1] = h;
        }
    }
    if (!prog.empty()) {
        ml.prog.resize(prog.size());
        upload(ml.prog.get(), prog.data(), prog.size() * sizeof(int64_t));
    }
    build_gather(ml, orow, rl);
    launches_.push_back(std::move(ml));
}

// Carried form. Stored level j's rows are rings of R = round64(c + T)
// samples, c being the look-back the launch reading j applies (its largest
// summed upper-branch shift) and W_max - 1 for searched rows; the stored
// levels are those with the fewest ring floats, launches spanning at most
// max_fuse_ levels and every plane level stored.
void Engine::carried_layout() {
    const uint32_t L = plan_.nlevels - 1;
    plane_setup();
    std::vector<std::vector<uint8_t>> searched(plan_.nlevels);
    for (uint32_t l = 0; l <= L; ++l) searched[l].assign(plan_.levels[l].size(), 0);
    for (const auto& p : plan_.planes)
        for (const int32_t r : p.level_row) searched[p.level][r] = 1;
    auto need_of = [&](uint32_t j, uint32_t l) {
        std::vector<int32_t> cur(plan_.levels[l].size(), 0), nxt;
        for (uint32_t m = l; m > j; --m) {
            const auto& v = plan_.levels[m];
            nxt.assign(plan_.levels[m - 1].size(), 0);
            for (std::size_t r = 0; r < v.size(); ++r) {
                int32_t& lo = nxt[v.chan_or_lo[r]];
                int32_t& hi = nxt[v.hi[r]];
                lo          = std::max(lo, cur[r]);
                hi          = std::max(hi, cur[r] + v.shift[r]);
            }
            cur.swap(nxt);
        }
        for (std::size_t r = 0; r < cur.size(); ++r)
            if (searched[j][r]) cur[r] = std::max<int32_t>(cur[r], static_cast<int32_t>(wm1_));
        return cur;
    };
    auto ring_floats = [&](const std::vector<int32_t>& need) {
        uint64_t n = 0;
        for (const int32_t c : need) n += round64(static_cast<uint64_t>(c) + tile_);
        return n;
    };
    std::vector<bool> must(plan_.nlevels, false);
    must[0] = must[L] = true;
    for (const auto& p : plan_.planes) must[p.level] = true;
    const uint32_t F = std::max<uint32_t>(1, std::min<uint32_t>(max_fuse_, 4));
    constexpr uint64_t kInf = ~uint64_t{0};
    std::vector<uint64_t> best(plan_.nlevels, kInf);
    std::vector<int32_t> prev(plan_.nlevels, -1);
    best[0] = 0; // level 0's rings are charged to the launch that reads them
    for (uint32_t l = 1; l <= L; ++l) {
        for (uint32_t j = l; j-- > 0 && l - j <= F;) {
            if (best[j] != kInf) {
                const uint64_t c = (j == 0 && x0_fused_) ? 0 : ring_floats(need_of(j, l));
                if (best[j] + c < best[l]) {
                    best[l] = best[j] + c;
                    prev[l] = static_cast<int32_t>(j);
                }
            }
            if (must[j]) break; // a plane level cannot be skipped
        }
    }
    carried_stored_.clear();
    for (int32_t l = static_cast<int32_t>(L); l > 0; l = prev[l]) {
        if (l < 0 || prev[l] < 0) throw std::logic_error("FDMT level fusion has no solution");
        carried_stored_.push_back(static_cast<uint32_t>(l));
    }
    carried_stored_.push_back(0);
    std::reverse(carried_stored_.begin(), carried_stored_.end());
    // Ring geometry, one region per stored level in one buffer.
    hist_.assign(plan_.nlevels, {});
    base_.assign(plan_.nlevels, {});
    ring_len_.assign(plan_.nlevels, {});
    level_len_.assign(plan_.nlevels, 0);
    for (uint32_t l = 0; l <= L; ++l) {
        hist_[l].assign(plan_.levels[l].size(), 0);
        base_[l].assign(plan_.levels[l].size(), 0);
        ring_len_[l].assign(plan_.levels[l].size(), 0);
    }
    uint64_t off = 0;
    for (std::size_t i = 0; i < carried_stored_.size(); ++i) {
        const uint32_t j = carried_stored_[i];
        hist_[j] = i + 1 < carried_stored_.size() ? need_of(j, carried_stored_[i + 1]) : need_of(j, j);
        if (j == 0 && x0_fused_) continue; // formed from the input: no rings
        for (std::size_t r = 0; r < hist_[j].size(); ++r) {
            const uint64_t R = round64(static_cast<uint64_t>(hist_[j][r]) + tile_);
            if (R >= (uint64_t{1} << 29))
                throw std::invalid_argument("FDMT ring longer than 2^29 samples; lower the time tile");
            base_[j][r]     = static_cast<int64_t>(off);
            ring_len_[j][r] = static_cast<int32_t>(R);
            off += R;
            level_len_[j] += R;
        }
    }
    rings_floats_ = off;
    // The input history (as build_init computes it): per channel, the
    // largest level-0 look-back read from the input plus its largest delay.
    const auto& l0 = plan_.levels[0];
    std::vector<uint64_t> hmax(plan_.nchans_fdmt, 0), dmax(plan_.nchans_fdmt, 0);
    for (std::size_t r = 0; r < l0.size(); ++r) {
        const int32_t c = l0.chan_or_lo[r];
        if (x0_fused_) hmax[c] = std::max<uint64_t>(hmax[c], static_cast<uint64_t>(hist_[0][r]));
        dmax[c] = std::max<uint64_t>(dmax[c], static_cast<uint64_t>(l0.delay[r]));
    }
    hist_in_ = 0;
    for (uint64_t c = 0; c < plan_.nchans_fdmt; ++c) hist_in_ = std::max(hist_in_, hmax[c] + dmax[c]);
    xstride_ = round64(hist_in_ + cap_);
}

void Engine::build_carried() {
    carried_layout();
    const auto& st = carried_stored_;
    level_source_.assign(plan_.nlevels, -2);
    level_buf_.assign(plan_.nlevels, -1);
    level_source_[0] = x0_fused_ ? -2 : -1;
    for (const uint32_t j : st)
        if (!(j == 0 && x0_fused_)) level_buf_[j] = 0; // every ring is in one buffer
    build_init();
    launches_.clear();
    for (std::size_t i = 1; i < st.size(); ++i) add_launch(st[i - 1], st[i]);
    build_searches();
}

uint64_t Engine::carried_need(uint64_t tile) {
    tile_ = tile;
    carried_layout();
    const uint64_t nch = plan_.nchans_fdmt;
    return (nch * xstride_ + rings_floats_) * sizeof(float);
}

// Splits a launch's output rows into runs whose distinct input rows (with
// the spread of shifts their consumers apply) fit the LDS budget, and emits
// the staged program. prog holds the direct program [out, src_0..K-1, h].
void Engine::build_gather(MergeLaunch& ml, const std::vector<GatherRow>& orow,
                          const std::vector<std::pair<int32_t, int32_t>>& rl) const {
    const int K        = ml.radix;
    const uint32_t a   = ml.in_level;
    const auto& base_a = base_[a];
    const auto& hist_a = hist_[a];
    const auto& ring_a = ring_len_[a];
    // The zero channel's row (level 0, plans with pass-through nodes): its
    // leaves are never staged; they read a zero slot placed after the
    // block's staged rows (one output row may hold it twice, at shifts far
    // apart).
    int32_t zero_row = -1;
    if (a == 0 && plan_.nchans_fdmt > plan_.cfg.nchans)
        for (std::size_t c = 0; c < plan_.levels[0].size(); ++c)
            if (plan_.levels[0].chan_or_lo[c] == static_cast<int32_t>(plan_.cfg.nchans))
                zero_row = static_cast<int32_t>(c);
    constexpr int32_t kZeroLeaf = -1; // record offset resolved to the zero slot below
    struct Lf {
        int32_t c;
        int32_t shmin, shmax;
    };
    std::vector<GatherBlock> gb;
    std::vector<GatherLeaf> gl;
    const std::size_t RS = 4 + static_cast<std::size_t>(K);
    std::vector<int32_t> rec(static_cast<std::size_t>(ml.nrows) * RS, 0);
    std::vector<Lf> cur, tent;
    // Input rows per block: the staging keeps each wave's rows in registers.
    const std::size_t max_leaves = gather_lds_ <= 4096 ? 16 : gather_lds_ <= 8192 ? 32
                                 : gather_lds_ <= 12288 ? 48 : 64;
    if (2 * static_cast<uint64_t>(gather_lds_) * sizeof(float) > 160 * 1024)
        throw std::invalid_argument("FROST_AMD_GATHER_LDS too large");
    ml.max_leaves = static_cast<uint32_t>(max_leaves);
    uint32_t first = 0, lds_max = 0;
    int32_t hmax = 0;
    // Level-0 rows formed from the input: rows of one channel share a
    // window, so a run is bounded by its channel count (16) and each
    // channel's rows must fit one grid (shift range + chunk within kX0Grid,
    // or kNodeGrid in the node form; delays within the window).
    const bool x0         = ml.from_input;
    // Node form: a block's rows span whole sub-bands over short chunks
    // (blocks are cut only where a sub-band's rows exceed the LDS budget
    // node_lds_ or kNodeRows).
    const bool node       = x0 && x0_node_;
    const int CL          = node ? kNodeChunk : kChunkG; // samples per staged chunk
    const auto& l0        = plan_.levels[0];
    // Channels per block: four windows per producer wave (run form), two per
    // producer wave (node form).
    const std::size_t max_chans = node ? kNodeChans : 16;
    if (x0) ml.max_leaves = static_cast<uint32_t>(max_chans);
    ml.node = node;
    auto x0_ok = [&](const std::vector<Lf>& v) {
        std::vector<std::array<int32_t, 4>> ch; // channel, min shmin, max shmax, max delay
        for (const auto& x : v) {
            const int32_t c = l0.chan_or_lo[x.c], d = l0.delay[x.c];
            bool found      = false;
            for (auto& e : ch)
                if (e[0] == c) {
                    e[1]  = std::min(e[1], x.shmin);
                    e[2]  = std::max(e[2], x.shmax);
                    e[3]  = std::max(e[3], d);
                    found = true;
                }
            if (!found) ch.push_back({c, x.shmin, x.shmax, d});
        }
        if (ch.size() > max_chans) return false;
        for (const auto& e : ch) {
            if (e[3] > kX0Dmax) return false;
            if (node ? e[2] - e[1] + kNodeChunk > kNodeGrid : e[2] - e[1] + kChunkG > kX0Grid)
                return false;
        }
        return true;
    };
    auto need = [&](const std::vector<Lf>& v, bool& ok) {
        uint64_t n = 0;
        ok         = true;
        for (const auto& x : v) {
            n += CL + (x.shmax - x.shmin) + (node ? 6 : 0); // node form: quad padding
            ok = ok && (x.shmax - x.shmin) <= kSpreadG;
        }
        if (x0) ok = ok && x0_ok(v);
        return n;
    };
    // Record order: the launch's rows, except that the node form groups the
    // rows of each sub-band (the rows whose lowest channel is the same; tree
    // and staggered sub-bands interleave in the level's order). A record
    // holds its row's output offset, so any order stores the same level.
    std::vector<uint32_t> ord(ml.nrows);
    for (uint32_t r = 0; r < ml.nrows; ++r) ord[r] = r;
    if (node) {
        std::vector<int32_t> key(ml.nrows, std::numeric_limits<int32_t>::max());
        for (uint32_t r = 0; r < ml.nrows; ++r)
            for (int k = 0; k < K; ++k) {
                const int32_t c = rl[static_cast<std::size_t>(r) * K + k].first;
                if (c != zero_row) key[r] = std::min(key[r], l0.chan_or_lo[c]);
            }
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t p, uint32_t q) { return key[p] < key[q]; });
    }
    std::vector<GatherChan> gc;
    // Sub-band (node) of each input row, its largest delay, and a union-find
    // over nodes that share a block (used for the dispatch order below).
    const auto& la     = plan_.levels[a];
    const int32_t nmax = la.node.empty() ? 0 : *std::max_element(la.node.begin(), la.node.end());
    std::vector<int32_t> ndmax(static_cast<std::size_t>(nmax) + 1, 0), par(ndmax.size());
    for (std::size_t c = 0; c < la.size(); ++c)
        ndmax[la.node[c]] = std::max(ndmax[la.node[c]], la.delay[c]);
    for (std::size_t i = 0; i < par.size(); ++i) par[i] = static_cast<int32_t>(i);
    auto find = [&](int32_t x) {
        while (par[x] != x) x = par[x] = par[par[x]];
        return x;
    };
    std::vector<double> bpos; // mean DM position of a block's input rows
    std::vector<int32_t> bnode;
    auto close = [&](uint32_t end) {
        {
            double pos = 0;
            const int32_t r0 = find(la.node[cur[0].c]);
            for (const auto& x : cur) {
                const int32_t nd = la.node[x.c];
                pos += static_cast<double>(la.delay[x.c]) / (ndmax[nd] + 1.0);
                par[find(nd)] = r0;
            }
            bpos.push_back(pos / static_cast<double>(cur.size()));
            bnode.push_back(la.node[cur[0].c]);
        }
        GatherBlock b{};
        b.leaf_first = static_cast<int32_t>(gl.size());
        b.nleaves    = static_cast<int32_t>(cur.size());
        if (x0 && !node && cur.size() > static_cast<std::size_t>(kX0LeafMax))
            throw std::logic_error("level-0 run holds more input rows than the kernel stages");
        b.row_first  = static_cast<int32_t>(first);
        b.nrows      = static_cast<int32_t>(end - first);
        b.hmax       = hmax;
        b.chan_first = static_cast<int32_t>(gc.size());
        b.nchan      = 0;
        if (x0) // channel groups, delays ascending
            std::sort(cur.begin(), cur.end(), [&](const Lf& p, const Lf& q) {
                return std::make_pair(l0.chan_or_lo[p.c], l0.delay[p.c]) <
                       std::make_pair(l0.chan_or_lo[q.c], l0.delay[q.c]);
            });
        std::vector<int32_t> slot(cur.size());
        int32_t off = 0;
        // Node form: each channel's largest shift (its grid origin); a row's
        // slot is padded to whole quads of grid positions, quad-aligned.
        std::vector<int32_t> cxs;
        if (node) {
            cxs.assign(plan_.nchans_fdmt + 1, std::numeric_limits<int32_t>::min());
            for (const auto& x : cur) {
                int32_t& v = cxs[l0.chan_or_lo[x.c]];
                v          = std::max(v, x.shmax);
            }
        }
        for (std::size_t i = 0; i < cur.size(); ++i) {
            const auto& x = cur[i];
            GatherLeaf f{};
            // Linear rows: offset of sample t0 - shmax; rings: the ring start.
            f.src   = carried_ ? base_a[x.c] : base_a[x.c] + hist_a[x.c] - x.shmax;
            f.slot  = off;
            f.len   = CL + (x.shmax - x.shmin);
            f.shmax = x.shmax;
            f.hc    = hist_a[x.c];
            f.delay = x0 ? l0.delay[x.c] : 0;
            f.ring  = carried_ && !x0 ? ring_a[x.c] : 0;
            int32_t area = f.len;
            if (node) {
                const int32_t plo = cxs[l0.chan_or_lo[x.c]] - x.shmax; // grid position of element 0
                const int32_t alo = plo & ~3, ahi = (plo + f.len + 3) & ~3;
                f.slot            = off + (plo - alo);
                area              = ahi - alo;
            }
            slot[i] = f.slot;
            off += area;
            if (x0) {
                const int32_t c = l0.chan_or_lo[x.c];
                if (b.nchan == 0 || gc.back().xoff != static_cast<int64_t>(c) * xstride_) {
                    GatherChan g{};
                    g.xoff       = static_cast<int64_t>(c) * xstride_;
                    g.xs         = x.shmax;
                    g.xmin       = x.shmin;
                    g.dmax       = f.delay;
                    g.leaf_first = static_cast<int32_t>(gl.size());
                    g.nleaf      = 0;
                    gc.push_back(g);
                    ++b.nchan;
                }
                auto& g = gc.back();
                g.xs    = std::max(g.xs, x.shmax);
                g.xmin  = std::min(g.xmin, x.shmin);
                g.dmax  = std::max(g.dmax, f.delay);
                g.dmask |= 1U << f.delay;
                ++g.nleaf;
            }
            gl.push_back(f);
        }
        lds_max = std::max<uint32_t>(lds_max, static_cast<uint32_t>(off));
        for (uint32_t r = first; r < end; ++r)
            for (int k = 0; k < K; ++k) {
                const auto [c, sh] = rl[static_cast<std::size_t>(ord[r]) * K + k];
                if (c == zero_row) {
                    rec[r * RS + 4 + k] = kZeroLeaf;
                    continue;
                }
                std::size_t i      = 0;
                while (cur[i].c != c) ++i;
                rec[r * RS + 4 + k] = slot[i] + (cur[i].shmax - sh);
            }
        gb.push_back(b);
    };
    for (uint32_t r = 0; r < ml.nrows; ++r) {
        const std::size_t sr = ord[r]; // the row whose record is r
        const int32_t h    = orow[sr].h;
        const int64_t o    = orow[sr].out;
        rec[r * RS + 0]    = static_cast<int32_t>(static_cast<uint32_t>(o & 0xffffffff));
        rec[r * RS + 1]    = static_cast<int32_t>(o >> 32);
        rec[r * RS + 2]    = h;
        rec[r * RS + 3]    = orow[sr].ring;
        tent               = cur;
        for (int k = 0; k < K; ++k) {
            const auto [c, sh32] = rl[sr * K + k];
            const int64_t sh     = sh32;
            if (sh < 0 || sh > std::numeric_limits<int32_t>::max() / 2 || h + sh > hist_a[c])
                throw std::logic_error("FDMT gather shift out of range");
            if (c == zero_row) continue;
            bool found = false;
            for (auto& x : tent)
                if (x.c == c) {
                    x.shmin = std::min(x.shmin, static_cast<int32_t>(sh));
                    x.shmax = std::max(x.shmax, static_cast<int32_t>(sh));
                    found   = true;
                }
            if (!found) tent.push_back({c, static_cast<int32_t>(sh), static_cast<int32_t>(sh)});
        }
        bool ok          = true;
        const uint64_t n = need(tent, ok);
        if (r > first && (!ok || n > (node ? node_lds_ : gather_lds_) ||
                          (!x0 && tent.size() > max_leaves) ||
                          r - first >= static_cast<uint32_t>(node ? kNodeRows : kGatherRowsMax))) {
            close(r);
            first = r;
            hmax  = 0;
            cur.clear();
            for (int k = 0; k < K; ++k) {
                const auto [c, sh] = rl[sr * K + k];
                if (c == zero_row) continue;
                bool found         = false;
                for (auto& x : cur)
                    if (x.c == c) {
                        x.shmin = std::min(x.shmin, sh);
                        x.shmax = std::max(x.shmax, sh);
                        found   = true;
                    }
                if (!found) cur.push_back({c, sh, sh});
            }
        } else {
            cur.swap(tent);
        }
        hmax = std::max(hmax, h);
    }
    if (ml.nrows > first) close(ml.nrows);
    {
        // Dispatch order: each XCD sweeps the block table in order and its L2
        // keeps the input rows of the blocks in flight. Blocks whose input
        // sub-bands overlap (tree and staggered sub-band rows of one level)
        // are interleaved by DM position, so that an input row both read is
        // still cached at the second read; groups of blocks with disjoint
        // inputs keep their order (stable sort).
        std::vector<int32_t> gfirst(par.size(), std::numeric_limits<int32_t>::max());
        for (std::size_t i = 0; i < gb.size(); ++i) {
            auto& f = gfirst[find(bnode[i])];
            f       = std::min(f, static_cast<int32_t>(i));
        }
        std::vector<uint32_t> order(gb.size());
        for (std::size_t i = 0; i < order.size(); ++i) order[i] = static_cast<uint32_t>(i);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t p, uint32_t q) {
            const int32_t fp = gfirst[find(bnode[p])], fq = gfirst[find(bnode[q])];
            return fp != fq ? fp < fq : bpos[p] < bpos[q];
        });
        std::vector<GatherBlock> sorted(gb.size());
        for (std::size_t i = 0; i < order.size(); ++i) sorted[i] = gb[order[i]];
        gb.swap(sorted);
    }
    ml.nblocks    = static_cast<uint32_t>(gb.size());
    ml.lds_floats = lds_max;
    ml.zero_slot  = -1;
    if (zero_row >= 0) {
        // One chunk of zeros after the largest block's staged rows.
        ml.zero_slot = static_cast<int32_t>((lds_max + 63) & ~63U);
        ml.lds_floats = static_cast<uint32_t>(ml.zero_slot) + CL;
        for (std::size_t r = 0; r < ml.nrows; ++r)
            for (int k = 0; k < K; ++k) {
                int32_t& v = rec[r * RS + 4 + static_cast<std::size_t>(k)];
                if (v == kZeroLeaf) v = ml.zero_slot;
            }
    }
    ml.blocks.resize(gb.size());
    ml.leaves.resize(gl.size());
    ml.rows.resize(rec.size());
    if (x0) {
        ml.chans.resize(gc.size());
        upload(ml.chans.get(), gc.data(), gc.size() * sizeof(GatherChan));
    }
    upload(ml.blocks.get(), gb.data(), gb.size() * sizeof(GatherBlock));
    upload(ml.leaves.get(), gl.data(), gl.size() * sizeof(GatherLeaf));
    upload(ml.rows.get(), rec.data(), rec.size() * 4);
}

void Engine::compute_inv(const std::vector<uint8_t>& mask, std::vector<float>& out) const {
    const std::size_t nw = widths_.size();
    // Distinct init lengths -> per-width window variance.
    std::vector<double> vtab((static_cast<std::size_t>(dmax0_) + 2) * nw, 0.0);
    for (int32_t a = 1; a <= dmax0_ + 1; ++a)
        for (std::size_t w = 0; w < nw; ++w)
            vtab[a * nw + w] = cfg_.use_trapezoid
                                   ? trapezoid_window_variance(a, widths_[w])
                                   : boxcar_window_variance(a, widths_[w]);
    const auto& l0 = plan_.levels[0];
    std::vector<double> prev(l0.size() * nw), cur;
    for (std::size_t r = 0; r < l0.size(); ++r)
        for (std::size_t w = 0; w < nw; ++w)
            prev[r * nw + w] = static_cast<std::size_t>(l0.chan_or_lo[r]) < mask.size() &&
                                       mask[l0.chan_or_lo[r]]
                                   ? vtab[(l0.delay[r] + 1) * nw + w]
                                   : 0.0; // masked, or the zero channel
    out.assign(static_cast<std::size_t>(nprows_) * nw, 0.0f);
    auto emit_level = [&](uint32_t l, const std::vector<double>& var) {
        for (std::size_t p = 0; p < plan_.planes.size(); ++p) {
            const auto& pd = plan_.planes[p];
            if (pd.level != l) continue;
            for (std::size_t i = 0; i < pd.delay.size(); ++i) {
                const std::size_t r = static_cast<std::size_t>(pd.level_row[i]);
                for (std::size_t w = 0; w < nw; ++w) {
                    const double v = var[r * nw + w];
                    out[(plane_first_[p] + i) * nw + w] =
                        v > 0.0 ? static_cast<float>(1.0 / std::sqrt(v)) : 0.0f;
                }
            }
        }
    };
    emit_level(0, prev);
    for (uint32_t l = 1; l < plan_.nlevels; ++l) {
        const auto& lv = plan_.levels[l];
        cur.assign(lv.size() * nw, 0.0);
        for (std::size_t r = 0; r < lv.size(); ++r) {
            const std::size_t lo = static_cast<std::size_t>(lv.chan_or_lo[r]);
            const std::size_t hi = static_cast<std::size_t>(lv.hi[r]);
            for (std::size_t w = 0; w < nw; ++w)
                cur[r * nw + w] = prev[lo * nw + w] + prev[hi * nw + w];
        }
        emit_level(l, cur);
        prev.swap(cur);
    }
}

std::vector<float> Engine::inv_sigma(uint32_t plane) const {
    if (plane >= plan_.planes.size()) throw std::invalid_argument("plane out of range");
    const std::size_t nw = widths_.size();
    const std::size_t b  = plane_first_[plane] * nw;
    const std::size_t n  = plan_.planes[plane].delay.size() * nw;
    if (inv_cur_host_.empty()) {
        std::vector<float> tmp;
        compute_inv(ctor_mask_, tmp);
        return std::vector<float>(tmp.begin() + b, tmp.begin() + b + n);
    }
    return std::vector<float>(inv_cur_host_.begin() + b, inv_cur_host_.begin() + b + n);
}

hipEvent_t Engine::sync_event(hipStream_t s) {
    if (sync_used_ == sync_events_.size()) {
        hipEvent_t e;
        FA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        sync_events_.push_back(e);
    }
    hipEvent_t e = sync_events_[sync_used_++];
    FA_HIP(hipEventRecord(e, s));
    return e;
}

void Engine::record_pair(int level, hipStream_t s, bool start) {
    if (capturing_) return; // captured gulps are timed as a whole
    if (start) {
        while (events_.size() < events_used_ + 2) {
            hipEvent_t e;
            FA_HIP(hipEventCreate(&e));
            events_.push_back(e);
            event_level_.push_back(0);
        }
        event_level_[events_used_] = level;
        FA_HIP(hipEventRecord(events_[events_used_], s));
    } else {
        FA_HIP(hipEventRecord(events_[events_used_ + 1], s));
        events_used_ += 2;
    }
}

void Engine::execute(const float* wf, uint64_t stride, uint64_t offset, uint64_t nsamps,
                     const uint8_t* kill_mask, hipStream_t s) {
    require(wf != nullptr, "waterfall must not be null");
    require(nsamps >= 1 && nsamps <= cap_, "nsamps must be in [1, block_nsamps]");
    require(stride >= offset + nsamps, "waterfall stride too small");
    require(!pending_collect_,
            "the previous gulp's candidates were not collected (call copy_result first)");
    require(x_ptr_ != nullptr, "no shared arena is bound to the engine");
    DeviceGuard dg(device_);
    FA_HIP(hipMemcpy2DAsync(input_columns(), xstride_ * 4, wf + offset, stride * 4, nsamps * 4,
                            plan_.cfg.nchans, hipMemcpyDeviceToDevice, s));
    execute_prepared(nsamps, kill_mask, s);
}

void Engine::execute_prepared(uint64_t N, const uint8_t* kill_mask, hipStream_t s) {
    prepare_gulp(N, kill_mask, s);
    DeviceGuard dg(device_);
    FA_HIP(hipEventRecord(ev_begin_, s));
    if (graph_mode_) {
        FA_HIP(hipGraphLaunch(gulp_graph(N, s), s));
    } else {
        enqueue_gulp(N, s);
    }
    FA_HIP(hipEventRecord(ev_end_, s));
    finish_gulp(N);
}

void Engine::prepare_gulp(uint64_t N, const uint8_t* kill_mask, hipStream_t s) {
    require(N >= 1 && N <= cap_, "nsamps must be in [1, block_nsamps]");
    require(s != nullptr, "an explicit non-null stream is required");
    require(!pending_collect_,
            "the previous gulp's candidates were not collected (call copy_result first)");
    require(x_ptr_ != nullptr, "no shared arena is bound to the engine");
    DeviceGuard dg(device_);
    const uint64_t nch = plan_.cfg.nchans;
    const std::vector<uint8_t> m =
        kill_mask ? std::vector<uint8_t>(kill_mask, kill_mask + nch) : ctor_mask_;
    const bool first = processed_ == 0;
    // Noise tables for this gulp's mask, and in inv_or_ the tables the
    // windows reaching into the previous gulp use: those of the OR of the two
    // masks after a change, else the current ones.
    if (inv_cur_host_.empty() || m != tables_mask_) {
        compute_inv(m, inv_cur_host_);
        tables_mask_ = m;
        FA_HIP(hipMemcpyAsync(inv_cur_.get(), inv_cur_host_.data(), inv_cur_host_.size() * 4,
                              hipMemcpyHostToDevice, s));
        floor_cur_host_.resize(inv_cur_host_.size());
        for (std::size_t i = 0; i < inv_cur_host_.size(); ++i)
            floor_cur_host_[i] = detect_floor(inv_cur_host_[i], cfg_.detect_thresh);
        FA_HIP(hipMemcpyAsync(floor_cur_.get(), floor_cur_host_.data(), floor_cur_host_.size() * 4,
                              hipMemcpyHostToDevice, s));
        or_is_cur_ = false;
    }
    const bool use_or = !first && m != prev_mask_;
    if (use_or) {
        std::vector<uint8_t> mor(nch);
        for (uint64_t c = 0; c < nch; ++c) mor[c] = static_cast<uint8_t>(m[c] | prev_mask_[c]);
        compute_inv(mor, inv_or_host_);
        FA_HIP(hipMemcpyAsync(inv_or_.get(), inv_or_host_.data(), inv_or_host_.size() * 4,
                              hipMemcpyHostToDevice, s));
        or_is_cur_ = false;
    } else if (!or_is_cur_) {
        FA_HIP(hipMemcpyAsync(inv_or_.get(), inv_cur_.get(), inv_cur_.bytes(),
                              hipMemcpyDeviceToDevice, s));
        or_is_cur_ = true;
    }
    if (keep_planes_ && keep_.size() < static_cast<uint64_t>(nprows_) * cap_) {
        keep_.reserve(static_cast<uint64_t>(nprows_) * cap_);
        drop_graphs(); // the plane-row store moved
    }
    // The gulp's absolute start sample, read by the kernels from device
    // memory (so a captured gulp replays for any gulp).
    set_u64_kernel<<<1, 64, 0, s>>>(reinterpret_cast<unsigned long long*>(params_.get()),
                                    processed_, 0);
    FA_HIP(hipGetLastError());
    pending_mask_ = m;
}

void Engine::finish_gulp(uint64_t N) {
    before_mask_ = prev_mask_;
    last_mask_   = pending_mask_;
    prev_mask_   = pending_mask_;
    processed_ += N;
    last_n_          = N;
    pending_collect_ = true;
}

// The gulp's device work, in stream order on s (and the search stream), with
// no host synchronisation, allocation or memset: it can be captured.
void Engine::enqueue_gulp(uint64_t N, hipStream_t s, bool captured) {
    DeviceGuard dg(device_);
    struct CaptureFlag {
        bool& f;
        bool prev;
        ~CaptureFlag() { f = prev; }
    } cf{capturing_, capturing_};
    capturing_ = capturing_ || captured;
    const uint64_t nch   = plan_.cfg.nchans;
    const std::size_t nw = widths_.size();
    if (shared_ && hist_in_ > 0)
        FA_HIP(hipMemcpy2DAsync(x_ptr_, xstride_ * 4, hist_store_.get(), hist_in_ * 4,
                                hist_in_ * 4, nch, hipMemcpyDeviceToDevice, s));
    if (shared_ && plan_.nchans_fdmt > nch) {
        // The zero channel's row of the shared arena (other engines write
        // there between this engine's gulps).
        zero_floats_kernel<<<ceil_div(xstride_, 256), 256, 0, s>>>(x_ptr_ + nch * xstride_, xstride_);
        FA_HIP(hipGetLastError());
    }
    // Open runs before this gulp: a re-run after a candidate overflow starts
    // from them.
    FA_HIP(hipMemcpyAsync(open_saved_.get(), open_.get(), open_.bytes(), hipMemcpyDeviceToDevice,
                          s));
    events_used_ = 0;
    sync_used_   = 0;
    // Searches still reading each level buffer (their last event; carried:
    // each level's rings), and the last search of the call.
    std::vector<hipEvent_t> reading(plan_.nlevels, nullptr);
    int flip    = 0; // this tile's buffer swap (flip_: every other tile)
    auto buf_id = [&](uint32_t l) { return carried_ ? static_cast<int>(l) : level_buf_[l] ^ flip; };
    hipEvent_t last_search = nullptr;
    const int gchunks      = gather_chunks();
    begin_gulp_kernel<<<1, 64, 0, s>>>(reinterpret_cast<unsigned long long*>(counters_.get()),
                                       params_.get(), N);
    const uint32_t L = plan_.nlevels - 1;
    for (uint64_t t0 = 0; t0 < N; t0 += tile_) {
        const int tl = static_cast<int>(std::min<uint64_t>(tile_, N - t0));
        flip         = flip_ ? static_cast<int>((t0 / tile_) & 1) : 0;
        for (uint32_t l = 0; l <= L; ++l) {
            if (level_source_[l] == -2) continue; // computed inside a fused launch
            float* out = buf_ptr_[level_buf_[l] ^ flip];
            if (hipEvent_t& r = reading[buf_id(l)]) {
                FA_HIP(hipStreamWaitEvent(s, r, 0)); // searches of the old contents done
                r = nullptr;
            }
            record_pair(static_cast<int>(l), s, true);
            if (l == 0) {
                int hcmax = 0; // computed look-back (none for rings)
                if (!carried_)
                    for (const int32_t h : hist_[0]) hcmax = std::max(hcmax, h);
                const unsigned gx = ceil_div(tl + hcmax + 64, kBlock * kInitPer);
                const uint64_t nx = plan_.nchans_fdmt;
                const unsigned gy = static_cast<unsigned>(std::min<uint64_t>(nx, 65535));
                const size_t lds  = (kBlock * kInitPer + dmax0_) * sizeof(float);
                fdmt_init_kernel<<<dim3(gx, gy), kBlock, lds, s>>>(
                    x_ptr_, xstride_, static_cast<int64_t>(hist_in_ + t0), init_chan_.get(),
                    init_rows_.get(), out, tl, static_cast<int>(nx),
                    reinterpret_cast<const uint64_t*>(params_.get()), static_cast<int64_t>(t0),
                    carried_ ? 1 : 0);
            } else {
                const MergeLaunch* ml = nullptr;
                for (const auto& mlc : launches_)
                    if (mlc.out_level == l) ml = &mlc;
                if (!ml) throw std::logic_error("FDMT level without a launch");
                const float* in = ml->from_input ? nullptr : buf_ptr_[ml->in_buf ^ flip];
                // Input samples of this tile: columns hist_in_ + t0 + tau,
                // tau in [xlo, xhi] inside the row.
                const float* x0 = x_ptr_ + hist_in_ + t0;
                const int xlo   = -static_cast<int>(hist_in_ + t0);
                const int xhi   = static_cast<int>(xstride_) - static_cast<int>(hist_in_ + t0) - 1;
                const auto* gabs_p = reinterpret_cast<const uint64_t*>(params_.get());
                if (staged_merge_) {
                    switch (ml->radix) {
                    case 2: launch_gather<2>(*ml, in, out, tl, x0, xlo, xhi, gchunks, carried_, gabs_p, static_cast<int64_t>(t0), s); break;
                    case 4: launch_gather<4>(*ml, in, out, tl, x0, xlo, xhi, gchunks, carried_, gabs_p, static_cast<int64_t>(t0), s); break;
                    case 8: launch_gather<8>(*ml, in, out, tl, x0, xlo, xhi, gchunks, carried_, gabs_p, static_cast<int64_t>(t0), s); break;
                    case 16: launch_gather<16>(*ml, in, out, tl, x0, xlo, xhi, gchunks, carried_, gabs_p, static_cast<int64_t>(t0), s); break;
                    default: throw std::logic_error("FDMT launch radix must be 2 to 16");
                    }
                } else {
                    switch (ml->radix) {
                    case 2: launch_merge<2>(in, out, ml->prog.get(), ml->nrows, tl, ml->hmax, s); break;
                    case 4: launch_merge<4>(in, out, ml->prog.get(), ml->nrows, tl, ml->hmax, s); break;
                    case 8: launch_merge<8>(in, out, ml->prog.get(), ml->nrows, tl, ml->hmax, s); break;
                    default: throw std::logic_error("direct FDMT merge supports radix 2 to 8");
                    }
                }
            }
            record_pair(static_cast<int>(l), s, false);
            if (search_count_[l]) {
                // The search of this level follows its launch on the search
                // stream (in level and tile order there, which the runs'
                // state across tiles needs).
                FA_HIP(hipStreamWaitEvent(search_stream_, sync_event(s), 0));
                record_pair(-1, search_stream_, true);
                launch_search(l, out, tl, t0, search_stream_);
                record_pair(-1, search_stream_, false);
                last_search = sync_event(search_stream_);
                reading[buf_id(l)] = last_search;
            }
        }
    }
    if (last_search) FA_HIP(hipStreamWaitEvent(s, last_search, 0)); // every search done
    // Frontier: earliest start of an open run, else the next sample.
    const uint64_t nopen = static_cast<uint64_t>(nprows_) * nw;
    frontier_kernel<<<ceil_div(nopen, 256), 256, 0, s>>>(
        open_.get(), nopen, reinterpret_cast<unsigned long long*>(counters_.get()) + 1);
    FA_HIP(hipGetLastError());
}

hipGraphExec_t Engine::gulp_graph(uint64_t N, hipStream_t s) {
    for (const auto& g : graphs_)
        if (g.arena == x_ptr_ && g.n == N) return g.exec;
    // Captured on its own stream (forked from and joined to s) so the
    // capture holds only this gulp's work. Events the enqueue records are
    // created by an eager gulp first, so the capture creates none.
    if (sync_events_.size() < 4 * plan_.nlevels * (N / tile_ + 2)) {
        capturing_ = false;
        warm_events(N);
    }
    if (!capture_stream_) FA_HIP(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking));
    FA_HIP(hipStreamSynchronize(s));
    hipGraph_t graph = nullptr;
    capturing_       = true;
    FA_HIP(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_gulp(N, capture_stream_, true);
    } catch (...) {
        (void)hipStreamEndCapture(capture_stream_, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        capturing_ = false;
        throw;
    }
    FA_HIP(hipStreamEndCapture(capture_stream_, &graph));
    capturing_ = false;
    hipGraphExec_t exec = nullptr;
    FA_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    graphs_.push_back({x_ptr_, N, graph, exec});
    return exec;
}

void Engine::warm_events(uint64_t N) {
    // Enough timing and ordering events for a gulp of N samples.
    const std::size_t need = 4 * plan_.nlevels * (N / tile_ + 2);
    while (sync_events_.size() < need) {
        hipEvent_t e;
        FA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        sync_events_.push_back(e);
    }
}

void Engine::drop_graphs() {
    ++graph_version_;
    for (auto& g : graphs_) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    graphs_.clear();
}

void Engine::set_graph_mode(bool on) { graph_mode_ = on; }

void Engine::set_gather_chunks(int chunks) {
    require(chunks >= 0 && chunks <= 64, "gather chunks must be in [0, 64]");
    if (chunks != chunks_) {
        chunks_ = chunks;
        drop_graphs();
    }
}

int Engine::gather_chunks() const { return chunks_ ? chunks_ : default_gather_chunks(); }

void Engine::set_search_kernel(int kind) {
    require(kind >= 0 && kind <= 2, "search kernel must be 0 (automatic), 1 or 2");
    const bool reg = kind == 2 ? false : reg_allowed_;
    require(kind != 1 || reg_allowed_, "the register search needs power-of-two widths up to 256");
    if (!reg) {
        int lds_max = 0;
        if (!sizing())
            FA_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device_));
        require(sizing() || static_cast<uint64_t>(ring_floats_) * sizeof(float) + 4096 <=
                                static_cast<uint64_t>(lds_max),
                "width trials need more LDS than the device offers for the ring-ladder search");
    }
    if (reg != reg_search_) {
        reg_search_ = reg;
        drop_graphs();
    }
}

void Engine::launch_search(uint32_t l, const float* buf, int tl, uint64_t t0, hipStream_t ss) {
    const std::size_t nw = widths_.size();
    SearchArgs a{};
    a.buf      = buf;
    a.rows     = search_rows_[l].get();
    a.nrows    = static_cast<int>(search_count_[l]);
    a.inv_cur  = inv_cur_.get();
    a.inv_or   = inv_or_.get();
    a.floor_cur = floor_cur_.get();
    a.widths   = widths_dev_.get();
    a.nw       = static_cast<int>(nw);
    a.kmax     = static_cast<int>(kmax_);
    a.wm1      = static_cast<int>(wm1_);
    a.tile_len = tl;
    a.tile_t0  = static_cast<int64_t>(t0);
    a.gulp_abs_p = reinterpret_cast<const uint64_t*>(params_.get());
    a.thr      = cfg_.detect_thresh;
    a.mode     = cfg_.peak_mode;
    a.trap     = cfg_.use_trapezoid ? 1 : 0;
    a.open     = open_.get();
    a.cands    = cands_.get();
    a.counter  = reinterpret_cast<unsigned long long*>(counters_.get());
    a.cap      = cand_cap_;
    a.spill    = spill_dev_;
    a.spill_cap = spill_cap_;
    a.keep     = keep_planes_ ? keep_.get() : nullptr;
    a.keep_stride = cap_;
    if (reg_search_) {
        RegSearchArgs ra{};
        ra.a       = a;
        ra.ring    = carried_;
        ra.present = 0;
        for (int k = 0; k < 16; ++k) ra.widx[k] = -1;
        for (std::size_t w = 0; w < nw; ++w) {
            const int k = 31 - __builtin_clz(widths_[w]);
            ra.present |= 1U << k;
            ra.widx[k] = static_cast<int>(w);
        }
        const unsigned blocks = ceil_div(search_count_[l], kBlock / 64);
        (cfg_.use_trapezoid ? reg_launcher<true>(static_cast<int>(kmax_))
                            : reg_launcher<false>(static_cast<int>(kmax_)))(ra, blocks, ss);
        return;
    }
    RingSearchArgs rs{};
    rs.a = a;
    for (int b = 0; b < kRingLevels; ++b) {
        rs.off[b]  = b <= static_cast<int>(kmax_) ? ring_off_[b] : 0;
        rs.size[b] = b <= static_cast<int>(kmax_) ? ring_size_[b] : 1;
    }
    rs.u0         = -static_cast<int>((wm1_ + kRingChunk - 1) / kRingChunk * kRingChunk);
    rs.lds_floats = ring_floats_;
    if (carried_)
        search_ring_kernel<true><<<search_count_[l], kRingChunk, ring_floats_ * sizeof(float), ss>>>(rs);
    else
        search_ring_kernel<false><<<search_count_[l], kRingChunk, ring_floats_ * sizeof(float), ss>>>(rs);
}

// Carries the last H input samples of each channel to the front of the
// extended input (the next gulp's look-back).
void Engine::carry_history(hipStream_t s) {
    const uint64_t nch = plan_.cfg.nchans;
    const uint64_t N   = last_n_;
    if (hist_in_ == 0) return;
    if (shared_) {
        // Columns [N, N + H) of the extended input are the next history
        // (the old history's tail first when N < H).
        FA_HIP(hipMemcpy2DAsync(hist_store_.get(), hist_in_ * 4, x_ptr_ + N, xstride_ * 4,
                                hist_in_ * 4, nch, hipMemcpyDeviceToDevice, s));
        return;
    }
    if (N >= hist_in_) {
        FA_HIP(hipMemcpy2DAsync(x_ptr_, xstride_ * 4, x_ptr_ + N, xstride_ * 4, hist_in_ * 4, nch,
                                hipMemcpyDeviceToDevice, s));
    } else {
        scratch_.reserve(nch * hist_in_);
        const unsigned gy = static_cast<unsigned>(std::min<uint64_t>(nch, 65535));
        copy_cols_kernel<<<dim3(ceil_div(hist_in_, 256), gy), 256, 0, s>>>(
            x_ptr_ + N, xstride_, scratch_.get(), hist_in_, hist_in_, nch);
        copy_cols_kernel<<<dim3(ceil_div(hist_in_, 256), gy), 256, 0, s>>>(
            scratch_.get(), hist_in_, x_ptr_, xstride_, hist_in_, nch);
    }
    FA_HIP(hipGetLastError());
}

void Engine::timing_collect() {
    std::fill(level_seconds.begin(), level_seconds.end(), 0.0);
    std::fill(level_launches.begin(), level_launches.end(), 0.0);
    search_seconds = 0;
    for (size_t i = 0; i + 1 < events_used_; i += 2) {
        const double t = event_seconds(events_[i], events_[i + 1]);
        const int l    = event_level_[i];
        if (l < 0) {
            search_seconds += t;
        } else {
            level_seconds[l] += t;
            level_launches[l] += 1;
        }
    }
    total_seconds = event_seconds(ev_begin_, ev_end_);
}

namespace {
void finalize(std::vector<fa_raw_candidate>& res, const fa_boxcar_config& cfg, uint64_t& dg,
              uint64_t& dd) {
    auto key = [](const fa_raw_candidate& c) {
        return std::make_tuple(c.subband_level, c.subband_idx, c.dm_idx, c.filter_idx,
                               c.peak_sample);
    };
    std::sort(res.begin(), res.end(),
              [&](const fa_raw_candidate& a, const fa_raw_candidate& b) { return key(a) < key(b); });
    std::vector<fa_raw_candidate> kept;
    kept.reserve(std::min<std::size_t>(res.size(), cfg.max_candidates));
    uint64_t run = 0;
    for (std::size_t i = 0; i < res.size(); ++i) {
        const bool same = i > 0 && res[i].subband_level == res[i - 1].subband_level &&
                          res[i].subband_idx == res[i - 1].subband_idx &&
                          res[i].dm_idx == res[i - 1].dm_idx;
        run = same ? run + 1 : 0;
        if (cfg.max_cands_per_dm > 0 && run >= cfg.max_cands_per_dm) {
            ++dd;
            continue;
        }
        if (kept.size() >= cfg.max_candidates) {
            ++dg;
            continue;
        }
        kept.push_back(res[i]);
    }
    res.swap(kept);
}
} // namespace

void Engine::collect(hipStream_t s) {
    DeviceGuard dg(device_);
    FA_HIP(hipStreamSynchronize(s));
    if (!pending_collect_) return; // the last call's view stays valid
    const auto h0    = std::chrono::steady_clock::now();
    result_.clear();
    dropped_global_ = dropped_per_dm_ = 0;
    // Copies on the gulp's stream (a synchronous hipMemcpy would also wait
    // for work on other streams, such as the next gulp's ingest copy).
    uint64_t ctr[2];
    FA_HIP(hipMemcpyAsync(ctr, counters_.get(), sizeof(ctr), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (ctr[0] > cand_cap_ && !carried_) {
        // More records than the buffer holds: which ones were stored depends
        // on the order of the atomic appends. Grow the buffer and search the
        // gulp again from the runs open before it (its input is still in
        // place: the history carry waits for this collect).
        if (ctr[0] > kCandCapMax)
            throw std::runtime_error(
                "the search emitted " + std::to_string(ctr[0]) +
                " raw candidates in one gulp, more than the " + std::to_string(kCandCapMax) +
                " the device buffer may hold; raise detect_thresh");
        cand_cap_ = ctr[0];
        cands_.resize(cand_cap_);
        drop_graphs(); // captured gulps hold the old buffer
        FA_HIP(hipMemcpyAsync(open_.get(), open_saved_.get(), open_.bytes(),
                              hipMemcpyDeviceToDevice, s));
        processed_ -= last_n_;
        prev_mask_       = before_mask_;
        pending_collect_ = false;
        const std::vector<uint8_t> m = last_mask_;
        ++reruns_;
        execute_prepared(last_n_, m.data(), s);
        collect(s);
        return;
    }
    pending_collect_ = false;
    carry_history(s);
    frontier_             = ctr[1];
    raw_candidates_       = ctr[0];
    // Carried form: records past the device buffer are in the host spill;
    // past both, some are lost (reported below, after this gulp's rows).
    const uint64_t stored = std::min<uint64_t>(ctr[0], cand_cap_ + spill_cap_);
    const uint64_t on_dev = std::min<uint64_t>(stored, cand_cap_);
    recs_.resize(stored);
    if (on_dev) {
        FA_HIP(hipMemcpyAsync(recs_.data(), cands_.get(), on_dev * sizeof(CandRecord),
                              hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
    }
    if (stored > on_dev) std::memcpy(recs_.data() + on_dev, spill_, (stored - on_dev) * sizeof(CandRecord));
    const bool lost = ctr[0] > stored;
    if (carried_ && ctr[0] > cand_cap_) {
        // Make room on the device for the next gulps.
        cand_cap_ = std::min<uint64_t>(std::max<uint64_t>(ctr[0], 2 * cand_cap_), kCandCapMax);
        cands_.resize(cand_cap_);
        drop_graphs();
    }
    // Order of frost's caps: (subband level, subband idx, dm idx, filter idx,
    // peak). Planes are numbered by (depth, idx) and a plane row is
    // plane_first + dm idx, so (plane row, width, peak) is the same order;
    // it is unique per call, so the order is deterministic. Sort the compact
    // records, apply the caps, and convert only the rows kept.
    // Counting pass on the plane row, then a small sort inside each row.
    row_start_.assign(static_cast<std::size_t>(nprows_) + 1, 0);
    for (const auto& r : recs_) ++row_start_[r.prow + 1];
    for (std::size_t i = 1; i < row_start_.size(); ++i) row_start_[i] += row_start_[i - 1];
    sorted_.resize(recs_.size());
    {
        std::vector<uint64_t> fill(row_start_.begin(), row_start_.end() - 1);
        for (const auto& r : recs_) sorted_[fill[r.prow]++] = r;
    }
    for (std::size_t i = 0; i + 1 < row_start_.size(); ++i)
        if (row_start_[i + 1] - row_start_[i] > 1)
            std::sort(sorted_.begin() + static_cast<std::ptrdiff_t>(row_start_[i]),
                      sorted_.begin() + static_cast<std::ptrdiff_t>(row_start_[i + 1]),
                      [](const CandRecord& a, const CandRecord& b) {
                          return a.w != b.w ? a.w < b.w : a.peak < b.peak;
                      });
    recs_.swap(sorted_);
    result_.reserve(std::min<uint64_t>(stored, cfg_.max_candidates));
    uint64_t run = 0;
    for (std::size_t k = 0; k < recs_.size(); ++k) {
        const CandRecord& r = recs_[k];
        run                 = (k > 0 && r.prow == recs_[k - 1].prow) ? run + 1 : 0;
        if (cfg_.max_cands_per_dm > 0 && run >= cfg_.max_cands_per_dm) {
            ++dropped_per_dm_;
            continue;
        }
        if (result_.size() >= cfg_.max_candidates) {
            ++dropped_global_;
            continue;
        }
        const auto it    = std::upper_bound(plane_first_.begin(), plane_first_.end(), r.prow);
        const auto p     = static_cast<std::size_t>(it - plane_first_.begin()) - 1;
        const auto& pd   = plan_.planes[p];
        const uint32_t i = r.prow - plane_first_[p];
        fa_raw_candidate c{};
        c.peak_sample   = r.peak;
        c.begin_sample  = r.begin;
        c.end_sample    = r.end;
        c.snr           = r.snr;
        c.dm            = pd.dm[i];
        c.dm_idx        = i;
        c.filter_idx    = r.w;
        c.width         = widths_[r.w];
        c.chan_begin    = pd.c0;
        c.chan_end      = pd.c1;
        c.f_low_mhz     = pd.f_low;
        c.f_high_mhz    = pd.f_high;
        c.subband_level = static_cast<uint16_t>(pd.depth);
        c.subband_idx   = static_cast<uint16_t>(pd.idx);
        c.staggered     = static_cast<uint8_t>(pd.staggered);
        c.beam_idx      = 0;
        result_.push_back(c);
    }
    collect_seconds_ =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
    timing_collect();
    if (lost)
        throw std::runtime_error(
            "the search emitted " + std::to_string(ctr[0]) + " raw candidates in one gulp, more " +
            "than the " + std::to_string(stored) + " the carried FDMT state can hold; the " +
            "gulp's rows are incomplete (raise detect_thresh)");
}

void Engine::drain(hipStream_t s) {
    require(!pending_collect_,
            "the last gulp's candidates were not collected (call copy_result before drain)");
    DeviceGuard dg(device_);
    FA_HIP(hipStreamSynchronize(s));
    const std::size_t nw = widths_.size();
    std::vector<OpenRun> st(static_cast<std::size_t>(nprows_) * nw);
    FA_HIP(hipMemcpyAsync(st.data(), open_.get(), st.size() * sizeof(OpenRun),
                          hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    result_.clear();
    for (std::size_t g = 0; g < st.size(); ++g) {
        if (!st[g].open) continue;
        const uint32_t prow = static_cast<uint32_t>(g / nw);
        const uint32_t w    = cfg_.peak_mode == FA_PEAK_ACROSS_WIDTHS
                                  ? st[g].w
                                  : static_cast<uint32_t>(g % nw);
        const auto it = std::upper_bound(plane_first_.begin(), plane_first_.end(), prow);
        const auto p  = static_cast<std::size_t>(it - plane_first_.begin()) - 1;
        const auto& pd = plan_.planes[p];
        const uint32_t i = prow - plane_first_[p];
        fa_raw_candidate c{};
        c.peak_sample   = st[g].peak;
        c.begin_sample  = st[g].start;
        c.end_sample    = st[g].end;
        c.snr           = st[g].snr;
        c.dm            = pd.dm[i];
        c.dm_idx        = i;
        c.filter_idx    = w;
        c.width         = widths_[w];
        c.chan_begin    = pd.c0;
        c.chan_end      = pd.c1;
        c.f_low_mhz     = pd.f_low;
        c.f_high_mhz    = pd.f_high;
        c.subband_level = static_cast<uint16_t>(pd.depth);
        c.subband_idx   = static_cast<uint16_t>(pd.idx);
        c.staggered     = static_cast<uint8_t>(pd.staggered);
        result_.push_back(c);
    }
    zero_open_kernel<<<ceil_div(st.size(), 256), 256, 0, s>>>(open_.get(), st.size());
    FA_HIP(hipGetLastError());
    FA_HIP(hipStreamSynchronize(s));
    dropped_global_ = dropped_per_dm_ = 0;
    finalize(result_, cfg_, dropped_global_, dropped_per_dm_);
    frontier_ = processed_;
}

uint64_t Engine::memory_bytes() const {
    uint64_t tables = 0;
    for (const auto& ml : launches_)
        tables += ml.prog.bytes() + ml.blocks.bytes() + ml.leaves.bytes() + ml.chans.bytes() +
                  ml.rows.bytes();
    for (const auto& sr : search_rows_) tables += sr.bytes();
    return x_.bytes() + buf_[0].bytes() + buf_[1].bytes() + rings_.bytes() + hist_store_.bytes() +
           inv_cur_.bytes() + inv_or_.bytes() + floor_cur_.bytes() + open_.bytes() + open_saved_.bytes() + cands_.bytes() + keep_.bytes() + scratch_.bytes() +
           init_chan_.bytes() + init_rows_.bytes() + widths_dev_.bytes() + counters_.bytes() +
           tables;
}

std::vector<uint64_t> Engine::level_samples() const {
    // Linear rows: the tile and its recomputed look-back; carried rows: the
    // ring's tile and carried look-back (only the tile is computed).
    std::vector<uint64_t> out(plan_.nlevels, 0);
    for (uint32_t l = 0; l < plan_.nlevels; ++l)
        for (const int32_t h : hist_[l]) out[l] += tile_ + static_cast<uint64_t>(h);
    return out;
}

void Engine::keep_planes(bool keep) { keep_planes_ = keep; }

void Engine::plane_rows(uint32_t plane, float* out, uint64_t cap, hipStream_t s) const {
    require(keep_planes_ && keep_.size() > 0, "plane rows were not kept");
    require(plane < plan_.planes.size(), "plane out of range");
    const uint64_t n = plan_.planes[plane].delay.size();
    require(cap >= n * last_n_, "output too small");
    DeviceGuard dg(device_);
    if (s) FA_HIP(hipStreamSynchronize(s));
    else FA_HIP(hipDeviceSynchronize());
    FA_HIP(hipMemcpy2D(out, last_n_ * 4, keep_.get() + static_cast<uint64_t>(plane_first_[plane]) * cap_,
                       cap_ * 4, last_n_ * 4, n, hipMemcpyDeviceToHost));
}

} // namespace fa
