// canon.hpp -- canonical form of a triangulation on the device: the simplices with only real
// vertices as sorted vertex tuples in lexicographic order, and the unique edges between real
// vertices, sorted (the form the parity tests compare; the reference's check_delaunay walks the
// DashMaps instead, src/delaunay_tree.rs:512-541, :642-671).
//
// Building blocks: a stable least-significant-digit radix sort of 64-bit keys (8-bit digits,
// optional 32-bit payload) and an exclusive scan of int counts.  Each sort pass is three
// kernels: per-tile digit histograms (tile = 4,096 elements, one block), a scan of the
// histograms in digit-major order (tile offsets of every digit), and a stable scatter in which
// each wavefront ranks its 64 elements per digit with ballots.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vo {

constexpr int kRadixThreads = 256;
constexpr int kRadixTile = 4096;                         // 16 elements per thread
constexpr int kRadixWaves = kRadixThreads / 64;
constexpr int kRadixPerWave = kRadixTile / kRadixWaves;  // contiguous elements per wavefront

// exclusive scan inside 256-element blocks, block totals to bsum
__global__ void __launch_bounds__(256) k_scan_blocks(const int* in, int* out, int* bsum, int64_t n) {
    __shared__ int s[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int add = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}

__global__ void __launch_bounds__(256) k_scan_add(int* out, const int* boff, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] += boff[blockIdx.x];
}

// per-tile histogram of the digit at `shift`: hist[d * ntiles + tile]
__global__ void __launch_bounds__(kRadixThreads) k_radix_hist(const unsigned long long* keys, int64_t n, int shift,
                                                             int* hist, int ntiles) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRadixTile;
    for (int i = threadIdx.x; i < kRadixTile; i += kRadixThreads) {
        const int64_t j = base + i;
        if (j < n) atomicAdd(&h[(int)((keys[j] >> shift) & 255ull)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile: wavefront w owns elements [w * kRadixPerWave, (w + 1) * ...) of
// the tile; it counts its digits, the block turns the counts into per-wavefront offsets on top
// of the tile's offsets (off[d * ntiles + tile]), and each wavefront then places its elements
// in order, ranking equal digits among its 64 lanes with ballots.
__global__ void __launch_bounds__(kRadixThreads) k_radix_scatter(const unsigned long long* keys, const uint32_t* vals,
                                                                unsigned long long* okeys, uint32_t* ovals, int64_t n,
                                                                int shift, const int* off, int ntiles) {
    __shared__ int cnt[kRadixWaves][256];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int d = lane; d < 256; d += 64) cnt[w][d] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRadixTile + (int64_t)w * kRadixPerWave;
    for (int i = lane; i < kRadixPerWave; i += 64) {
        const int64_t j = base + i;
        if (j < n) atomicAdd(&cnt[w][(int)((keys[j] >> shift) & 255ull)], 1);
    }
    __syncthreads();
    // per-wavefront starting positions: tile offset of the digit + counts of earlier wavefronts
    for (int d = threadIdx.x; d < 256; d += kRadixThreads) {
        int acc = off[(int64_t)d * ntiles + blockIdx.x];
        for (int v = 0; v < kRadixWaves; ++v) {
            const int c = cnt[v][d];
            cnt[v][d] = acc;
            acc += c;
        }
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int i0 = 0; i0 < kRadixPerWave; i0 += 64) {
        const int64_t j = base + i0 + lane;
        const bool valid = j < n;
        unsigned long long k = 0;
        uint32_t v = 0;
        int d = 256;                                   // invalid lanes form a group of their own
        if (valid) {
            k = keys[j];
            if (vals) v = vals[j];
            d = (int)((k >> shift) & 255ull);
        }
        unsigned long long same = __ballot(true);
        for (int b = 0; b < 9; ++b) {
            const unsigned long long bb = __ballot((d >> b) & 1);
            same &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int rank = __popcll(same & below);
        const int start = valid ? cnt[w][d] : 0;
        if (valid) {
            const int pos = start + rank;
            okeys[pos] = k;
            if (vals) ovals[pos] = v;
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) cnt[w][d] = start + __popcll(same);
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void k_iota_u32(uint32_t* v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}

// Real simplices (every vertex a point: reference id >= first) as sorted tuples of point
// indices (id - first); flag[i] = 1 for them.  Rows of the output have stride 4 (2D: w = -1).
template <int M>
__global__ void k_canon_rows(const int32_t* simp, int64_t n, int first, int4* rows, int* flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int r[4] = {0, 0, 0, 0x7fffffff};
        bool real = true;
        for (int k = 0; k < M; ++k) {
            const int v = simp[i * M + k];
            real &= v >= first;
            r[k] = v - first;
        }
        // sorting network on M <= 4 values
        auto cs = [&](int a, int b) {
            const int lo = r[a] < r[b] ? r[a] : r[b], hi = r[a] < r[b] ? r[b] : r[a];
            r[a] = lo;
            r[b] = hi;
        };
        if (M == 4) { cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2); }
        else { cs(0, 1); cs(1, 2); cs(0, 1); }
        rows[i] = make_int4(r[0], r[1], r[2], M == 4 ? r[3] : -1);
        flag[i] = real;
    }
}

// compacted rows and their two 64-bit sort keys: lo = columns (2, 3) (2D: column 2), hi =
// columns (0, 1), each column `nb` bits wide
template <int M>
__global__ void k_canon_keys(const int4* rows, const int* flag, const int* pos, int64_t n, int nb, int4* crow,
                             unsigned long long* klo, unsigned long long* khi) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const int4 r = rows[i];
        const int64_t o = pos[i];
        crow[o] = r;
        klo[o] = M == 4 ? (((unsigned long long)r.z << nb) | (unsigned long long)r.w) : (unsigned long long)r.z;
        khi[o] = ((unsigned long long)r.x << nb) | (unsigned long long)r.y;
    }
}

// key of every row in the permutation's order: dst[i] = khi[perm[i]]
__global__ void k_gather_u64(const unsigned long long* src, const uint32_t* perm, unsigned long long* dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

// rows in the sorted order, M columns
template <int M>
__global__ void k_canon_out_rows(const int4* crow, const uint32_t* perm, int64_t n, int32_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int4 r = crow[perm[i]];
        out[i * M + 0] = r.x;
        out[i * M + 1] = r.y;
        out[i * M + 2] = r.z;
        if (M == 4) out[i * M + 3] = r.w;
    }
}

// every vertex pair of every simplex as a key (min << nb | max); pairs with a super vertex get
// the largest key, which sorts last
template <int M>
__global__ void k_canon_edge_keys(const int32_t* simp, int64_t n, int first, int nb, unsigned long long* keys) {
    constexpr int P = M * (M - 1) / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int v[4];
        for (int k = 0; k < M; ++k) v[k] = simp[i * M + k];
        int e = 0;
        for (int a = 0; a < M; ++a)
            for (int b = a + 1; b < M; ++b) {
                unsigned long long key = ~0ull;
                if (v[a] >= first && v[b] >= first) {
                    const int x = v[a] - first, y = v[b] - first;
                    key = ((unsigned long long)(x < y ? x : y) << nb) | (unsigned long long)(x < y ? y : x);
                }
                keys[i * P + e++] = key;
            }
    }
}

// first occurrence of each real edge key in the sorted list
__global__ void k_canon_edge_flags(const unsigned long long* keys, int64_t n, int* flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = keys[i] != ~0ull && (i == 0 || keys[i] != keys[i - 1]);
}

__global__ void k_canon_edge_out(const unsigned long long* keys, const int* flag, const int* pos, int64_t n, int nb,
                                 int32_t* out) {
    const unsigned long long mask = (1ull << nb) - 1ull;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const int64_t o = pos[i];
        out[2 * o] = (int32_t)(keys[i] >> nb);
        out[2 * o + 1] = (int32_t)(keys[i] & mask);
    }
}

}  // namespace vo
