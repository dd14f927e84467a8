// Stable LSD radix sort, see radix_sort.h.
#include "radix_sort.h"

namespace godar {

// Per-tile digit histogram, written digit-major: hist[d * ntiles + tile].
__global__ __launch_bounds__(RS_THREADS) void rs_hist_kernel(const uint32_t* __restrict__ keys,
                                                             int64_t n, int shift, int ntiles,
                                                             uint32_t* __restrict__ hist,
                                                             const uint32_t* err, const int32_t* run) {
    if ((err && *err) || (run && !*run)) return;
    __shared__ uint32_t h[256];
    int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int k = 0; k < RS_IPT; ++k) {
        int64_t p = base + (int64_t)k * RS_THREADS + t;
        if (p < n) atomicAdd(&h[(keys[p] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)t * ntiles + blockIdx.x] = h[t];
}

// Exclusive scan of each digit's row of per-tile counts (one workgroup per
// digit); the row total goes to dtot[digit].
__global__ __launch_bounds__(256) void rs_scan_kernel(uint32_t* __restrict__ hist, int ntiles,
                                                      uint32_t* __restrict__ dtot, const uint32_t* err,
                                                      const int32_t* go) {
    if ((err && *err) || (go && !*go)) return;
    __shared__ uint32_t part[256];
    const int t = threadIdx.x;
    uint32_t* row = hist + (int64_t)blockIdx.x * ntiles;
    const int per = (ntiles + 255) / 256;
    const int lo = t * per, hi = lo + per < ntiles ? lo + per : ntiles;
    uint32_t sum = 0;
    for (int k = lo; k < hi; ++k) sum += row[k];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (int k = lo; k < hi; ++k) {
        uint32_t v = row[k];
        row[k] = run;
        run += v;
    }
    if (t == 255) dtot[blockIdx.x] = part[255];
}

// Stable scatter.  Items of a tile are visited in rounds of 256 (round r
// holds items r*256 .. r*256+255 in thread order), so the output keeps the
// input order among equal digits.
__global__ __launch_bounds__(RS_THREADS) void rs_scatter_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, int ntiles, const uint32_t* __restrict__ hist,
    const uint32_t* __restrict__ dtot, const uint32_t* err, const int32_t* run) {
    if ((err && *err) || (run && !*run)) return;
    constexpr int NW = RS_THREADS / 64;
    __shared__ uint32_t base_off[256];
    __shared__ uint32_t wcount[NW][256];
    int t = threadIdx.x;
    int w = t >> 6, lane = t & 63;
    // digit bases: exclusive scan of the digit totals
    base_off[t] = dtot[t];
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = t >= off ? base_off[t - off] : 0u;
        __syncthreads();
        base_off[t] += v;
        __syncthreads();
    }
    uint32_t excl = t ? base_off[t - 1] : 0u;
    __syncthreads();
    base_off[t] = excl + hist[(int64_t)t * ntiles + blockIdx.x];
    const uint64_t lt_mask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_IPT; ++r) {
#pragma unroll
        for (int q = 0; q < NW; ++q) wcount[q][t] = 0;
        __syncthreads();
        int64_t p = base + (int64_t)r * RS_THREADS + t;
        bool valid = p < n;
        uint32_t key = valid ? kin[p] : 0u;
        uint32_t val = valid ? vin[p] : 0u;
        uint32_t d = (key >> shift) & 255u;
        // lanes of this wavefront holding the same digit (9 ballots incl. validity)
        uint64_t m = __ballot(valid);
        if (!valid) m = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            m &= bit ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(m & lt_mask);
        if (valid && (m & lt_mask) == 0) wcount[w][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pre = 0;
            for (int q = 0; q < w; ++q) pre += wcount[q][d];
            uint32_t pos = base_off[d] + pre + rank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) tot += wcount[q][t];
        base_off[t] += tot;
        __syncthreads();
    }
}

hipError_t radix_alloc(RadixScratch& s, int64_t n) {
    if (n <= s.capacity && s.keys_alt) return hipSuccess;
    radix_free(s);
    int64_t cap = n < 1024 ? 1024 : n;
    int tiles = (int)((cap + RS_TILE - 1) / RS_TILE);
    hipError_t e;
    if ((e = hipMalloc(&s.keys_alt, cap * sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = hipMalloc(&s.vals_alt, cap * sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = hipMalloc(&s.hist, (int64_t)256 * tiles * sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = hipMalloc(&s.dtot, 256 * sizeof(uint32_t))) != hipSuccess) return e;
    s.capacity = cap;
    s.max_tiles = tiles;
    return hipSuccess;
}

void radix_free(RadixScratch& s) {
    if (s.keys_alt) (void)hipFree(s.keys_alt);
    if (s.vals_alt) (void)hipFree(s.vals_alt);
    if (s.hist) (void)hipFree(s.hist);
    if (s.dtot) (void)hipFree(s.dtot);
    s = RadixScratch{};
}

hipError_t radix_sort_pairs(RadixScratch& s, uint32_t* keys, uint32_t* vals, int64_t n, int bits,
                            const uint32_t* err, const int32_t* run, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (n > s.capacity) return hipErrorInvalidValue;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    // an odd number of passes ends in the caller's arrays (the input is in
    // the scratch pair); an even count gets one more pass on the top digit,
    // which leaves a sequence already sorted by it as it is (stable)
    const int total = passes | 1;
    int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    uint32_t *ka = s.keys_alt, *va = s.vals_alt, *kb = keys, *vb = vals;
    for (int ps = 0; ps < total; ++ps) {
        const int shift = (ps < passes ? ps : passes - 1) * 8;
        hipLaunchKernelGGL(rs_hist_kernel, dim3(ntiles), dim3(RS_THREADS), 0, stream, ka, n, shift,
                           ntiles, s.hist, err, run);
        hipLaunchKernelGGL(rs_scan_kernel, dim3(256), dim3(256), 0, stream, s.hist, ntiles, s.dtot,
                           err, run);
        hipLaunchKernelGGL(rs_scatter_kernel, dim3(ntiles), dim3(RS_THREADS), 0, stream, ka, va, kb,
                           vb, n, shift, ntiles, s.hist, s.dtot, err, run);
        uint32_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    return hipGetLastError();
}

}  // namespace godar
