// sort.hip -- stable LSD radix sort of (64-bit key, 32-bit value) pairs, the
// Morton sort of the tree build.
//
// Eight-bit digits, three launches per digit: the digit histogram of every
// tile of 2048 keys, one exclusive scan per digit over the tiles (digit-major,
// so a digit's scan gives each tile's first output position within the
// digit), and the rank-and-scatter pass.  No device fills and no host round
// trips: the whole sort can be recorded in a HIP graph (the library sort it
// replaces clears its state with memset nodes).
//
// Ranking inside a tile is stable: wave w owns keys [512 w, 512 w + 512) of
// its tile and takes them in 8 rounds of 64 consecutive keys; in a round each
// lane finds the lanes holding its digit with 8 ballots (one per digit bit),
// its rank is the number of those below it plus the count of the digit in the
// wave's earlier rounds (a per-wave LDS histogram that the highest such lane
// advances).  After the rounds the per-wave counts become per-wave bases, the
// tile's keys are placed in LDS in digit order, and each thread writes
// consecutive LDS slots to their global positions, so every digit's run of the
// tile is written contiguously.  A stable sort is unique, so the permutation is
// the one any other stable sort of the same keys gives.
#include "internal.hpp"

namespace sphgpu {
namespace {

constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / kWave;
constexpr int kRsRounds = 8;
constexpr int kRsTile = kRsThreads * kRsRounds;  // 2048 keys
constexpr int kRsBins = 256;
static_assert(kRsThreads == kRsBins, "one thread per digit in the per-digit steps");

__device__ __forceinline__ int rs_digit(uint64_t k, int shift) { return (int) ((k >> shift) & (kRsBins - 1)); }

// position of round r's key of this lane inside the tile
__device__ __forceinline__ int rs_slot(int w, int r, int lane) { return w * (kRsRounds * kWave) + r * kWave + lane; }

// inclusive scan over a wave
__device__ __forceinline__ int rs_wave_scan(int v, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// per tile: the count of each digit, to bh[digit * nb + tile]
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(int64_t n, const uint64_t *key, int shift, int nb,
                                                        int32_t *bh) {
    __shared__ int hist[kRsWaves][kRsBins];
    const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
    for (int i = tid; i < kRsWaves * kRsBins; i += kRsThreads) (&hist[0][0])[i] = 0;
    __syncthreads();
    const int64_t tile0 = (int64_t) blockIdx.x * kRsTile;
    uint64_t k[kRsRounds];
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
        const int64_t i = tile0 + rs_slot(w, r, lane);
        k[r] = (i < n) ? key[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r)
        if (tile0 + rs_slot(w, r, lane) < n) atomicAdd(&hist[w][rs_digit(k[r], shift)], 1);
    __syncthreads();
    int c = 0;
#pragma unroll
    for (int v = 0; v < kRsWaves; ++v) c += hist[v][tid];
    bh[(int64_t) tid * nb + blockIdx.x] = c;
}

// one workgroup per digit: exclusive scan of the digit's tile counts in place,
// the digit's total to tot[digit]
__global__ __launch_bounds__(kRsThreads) void k_rs_scan(int nb, int32_t *bh, int32_t *tot) {
    __shared__ int wsum[kRsWaves];
    const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
    int32_t *row = bh + (int64_t) blockIdx.x * nb;
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += kRsThreads) {
        const int b = c0 + tid;
        const int v = (b < nb) ? row[b] : 0;
        const int incl = rs_wave_scan(v, lane);
        if (lane == kWave - 1) wsum[w] = incl;
        __syncthreads();
        int pre = carry, chunk = 0;
#pragma unroll
        for (int u = 0; u < kRsWaves; ++u) {
            pre += (u < w) ? wsum[u] : 0;
            chunk += wsum[u];
        }
        if (b < nb) row[b] = pre + incl - v;
        carry += chunk;
        __syncthreads();
    }
    if (tid == 0) tot[blockIdx.x] = carry;
}

// rank the tile's keys by digit (stable) and write them to their positions;
// vin == nullptr takes the key's input index as its value
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(int64_t n, const uint64_t *kin, const int32_t *vin,
                                                           uint64_t *kout, int32_t *vout, int shift, int nb,
                                                           const int32_t *bh, const int32_t *tot) {
    __shared__ uint64_t sk[kRsTile];
    __shared__ int32_t sv[kRsTile];
    __shared__ int hist[kRsWaves][kRsBins];  // per-wave counts, then per-wave bases
    __shared__ int tstart[kRsBins];          // first slot of each digit in the tile's LDS order
    __shared__ int gofs[kRsBins];            // global position of the tile's first key of each digit
    __shared__ int wsum[2][kRsWaves];
    const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
    for (int i = tid; i < kRsWaves * kRsBins; i += kRsThreads) (&hist[0][0])[i] = 0;
    __syncthreads();
    const int64_t tile0 = (int64_t) blockIdx.x * kRsTile;
    const uint64_t lt = (1ULL << lane) - 1;  // lanes below this one
    uint64_t k[kRsRounds];
    int32_t val[kRsRounds];
    int rank[kRsRounds], dig[kRsRounds];
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
        const int64_t i = tile0 + rs_slot(w, r, lane);
        const bool ok = i < n;
        k[r] = ok ? kin[i] : 0;
        val[r] = ok ? (vin ? vin[i] : (int32_t) i) : 0;
    }
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
        const bool ok = tile0 + rs_slot(w, r, lane) < n;
        const int d = rs_digit(k[r], shift);
        dig[r] = d;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        // the wave's count of digit d so far; the highest lane holding d
        // advances it (LDS reads and writes of a wave take effect in order)
        const int h = ok ? hist[w][d] : 0;
        rank[r] = h + __popcll(peers & lt);
        if (ok && (peers >> lane) == 1) hist[w][d] = h + __popcll(peers);
    }
    __syncthreads();
    {  // thread d: per-wave bases, the tile's digit starts and the global ones
        const int d = tid;
        int c = 0;
#pragma unroll
        for (int v = 0; v < kRsWaves; ++v) {
            const int t = hist[v][d];
            hist[v][d] = c;
            c += t;
        }
        const int g = tot[d];
        const int ic = rs_wave_scan(c, lane), ig = rs_wave_scan(g, lane);
        if (lane == kWave - 1) {
            wsum[0][w] = ic;
            wsum[1][w] = ig;
        }
        __syncthreads();
        int pc = 0, pg = 0;
#pragma unroll
        for (int u = 0; u < kRsWaves; ++u)
            if (u < w) {
                pc += wsum[0][u];
                pg += wsum[1][u];
            }
        tstart[d] = pc + ic - c;
        gofs[d] = pg + ig - g + bh[(int64_t) d * nb + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
        if (tile0 + rs_slot(w, r, lane) < n) {
            const int d = dig[r];
            const int pos = tstart[d] + hist[w][d] + rank[r];
            sk[pos] = k[r];
            sv[pos] = val[r];
        }
    }
    __syncthreads();
    const int tn = (int) ((n - tile0 < kRsTile) ? (n - tile0) : kRsTile);
    for (int sl = tid; sl < tn; sl += kRsThreads) {
        const uint64_t key = sk[sl];
        const int d = rs_digit(key, shift);
        const int64_t o = (int64_t) gofs[d] + (sl - tstart[d]);
        kout[o] = key;
        vout[o] = sv[sl];
    }
}

inline int nblk(int64_t n, int b) { return (int) ((n + b - 1) / b); }

}  // namespace

int radix_sort_tiles(int64_t n) { return nblk(n, kRsTile); }

void radix_sort_pairs(int64_t n, int keybits, uint64_t *k0, uint64_t *k1, int32_t *v0, int32_t *v1, int32_t *bh,
                      int32_t *tot, uint64_t **kres, int32_t **vres, hipStream_t s) {
    const int nb = radix_sort_tiles(n);
    uint64_t *kin = k0, *kout = k1;
    int32_t *vin = nullptr, *vout = v0, *vspare = v1;
    for (int shift = 0; shift < keybits; shift += 8) {
        if (n > 0) {
            hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kRsThreads), 0, s, n, kin, shift, nb, bh);
            SPH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_rs_scan, dim3(kRsBins), dim3(kRsThreads), 0, s, nb, bh, tot);
            SPH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kRsThreads), 0, s, n, kin, vin, kout, vout, shift, nb, bh,
                               tot);
            SPH_LAUNCH_CHECK();
        }
        std::swap(kin, kout);
        vin = vout;
        vout = vspare;
        vspare = vin;
    }
    *kres = kin;
    *vres = vin;
}

}  // namespace sphgpu
