// compact.hpp -- compaction of the append-only simplex records of one triangulation.
//
// Killed simplices stay in the record array (their forward pointers re-bucket stale
// references lazily), so after k insertions about 27k records hold 6.7k live simplices (3D),
// scattered over the whole array in creation order.  Every visit of a simplex then lands on a
// cold line and, at large n, a cold page.  Compaction moves the live records to a fresh array
// in spatial order (4096 bins of a grid over the bounding box, by record centroid), renumbers
// adjacency, conflict pointers of pending points and walk hints, and resets reservations.
// It runs between rounds, when no point holds a reservation or a queued predicate.
#pragma once
#include "kernels.hpp"

namespace vo {

constexpr int kCompactBins = 4096;           // 16^3 (3D) / 64^2 (2D) cells

// follow forward pointers from t to a live simplex
template <int D> __device__ __forceinline__ int chase_live(const Ctx& c, int t, unsigned* err) {
    for (int hops = 0; t >= 0; ++hops) {
        const int* v = v_ptr<D>(c, t, 0);
        if (v[0] >= 0) return t;
        if (hops > kChainLimit) {
            atomicOr(err, ERR_CHAIN);
            return -1;
        }
        t = v[1];
    }
    return t;
}

// conflict pointers of the listed points (the pending list) move to live simplices
template <int D> __global__ void k_chase_points(Ctx c, const int* list, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int p = list[i];
        const int t = c.cp[p];
        if (t >= 0) c.cp[p] = chase_live<D>(c, t, &c.st->err);
    }
}

// walk hints likewise (level 0 always holds a simplex)
template <int D> __global__ void k_chase_hints(Ctx c, int* h, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int t = h[i];
        if (t >= 0) h[i] = chase_live<D>(c, t, &c.st->err);
    }
}

// bin of a live record: the grid cell of its vertices' centroid (translated frame)
template <int D> __device__ __forceinline__ int record_bin(const Ctx& c, int t, const float* glo, float ginv) {
    typename VT<D>::F x[4];
    load_tx<D>(c, t, x);
    constexpr int bits = D == 3 ? 4 : 6;
    constexpr float g = (float)(1 << bits);
    float s[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < D + 1; ++k) {
        s[0] += x[k].x;
        s[1] += x[k].y;
        if constexpr (D == 3) s[2] += x[k].z;
    }
    int q[3] = {0, 0, 0};
    for (int j = 0; j < D; ++j) {
        const float u = (s[j] / (float)(D + 1) - glo[j]) * ginv * g;
        q[j] = (int)fminf(fmaxf(u, 0.0f), g - 1.0f);       // super-simplex corners clamp to the edge
    }
    int code = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int j = D - 1; j >= 0; --j) code = (code << 1) | ((q[j] >> b) & 1);
    return code;
}

// per-bin counts of live records (block histogram in LDS, one global atomic per bin and block)
template <int D> __global__ void __launch_bounds__(256) k_bin_count(Ctx c, int n, int* hist, float g0, float g1, float g2,
                                                                    float ginv) {
    __shared__ int s_h[kCompactBins];
    for (int i = threadIdx.x; i < kCompactBins; i += 256) s_h[i] = 0;
    __syncthreads();
    const float glo[3] = {g0, g1, g2};
    for (int t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256)
        if (*v_ptr<D>(c, t, 0) >= 0) atomicAdd(&s_h[record_bin<D>(c, t, glo, ginv)], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < kCompactBins; i += 256)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// new index of every live record: its bin's next free position (order inside a bin is free)
template <int D> __global__ void __launch_bounds__(256) k_bin_place(Ctx c, int n, int* cursor, int* map, float g0, float g1,
                                                                    float g2, float ginv) {
    const float glo[3] = {g0, g1, g2};
    for (int t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256)
        map[t] = *v_ptr<D>(c, t, 0) >= 0 ? atomicAdd(&cursor[record_bin<D>(c, t, glo, ginv)], 1) : -1;
}

// live records to their new places, adjacency renumbered, reservations at the largest key
template <int D> __global__ void __launch_bounds__(256) k_move(Ctx c, char* dst, const int* map, int n) {
    Ctx d = c;
    d.rec = dst;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) {
        const int nt = map[t];
        if (nt < 0) continue;
        st_v<D>(d, nt, ld_v<D>(c, t));
        for (int k = 0; k < D + 1; ++k) {
            const int nb = *a_ptr<D>(c, t, k);
            *a_ptr<D>(d, nt, k) = nb < 0 ? -1 : ((map[nb >> 2] << 2) | (nb & 3));
        }
        typename VT<D>::F x[4];
        load_tx<D>(c, t, x);
        store_tx<D>(d, nt, x);
        *kres_p<D>(d, nt) = ~0ull;
        *nres_p<D>(d, nt) = ~0ull;
    }
}

__global__ void k_remap_points(int* cp, const int* list, int n, const int* map) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int p = list[i];
        if (cp[p] >= 0) cp[p] = map[cp[p]];
    }
}

__global__ void k_remap_hints(int* h, int n, const int* map) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (h[i] >= 0) h[i] = map[h[i]];
}

}  // namespace vo
