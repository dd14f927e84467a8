// Stable LSD radix sort of (uint32 key, uint32 value) pairs on one HIP stream.
// 8-bit digits; one pass = per-tile digit histogram -> exclusive scan of the
// (digit, tile) table -> stable scatter with wavefront-ballot ranking.
// Used to bin floes into the cell list (replaces the coretran k-d tree).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace godar {

struct RadixScratch {
    uint32_t* keys_alt = nullptr;
    uint32_t* vals_alt = nullptr;
    uint32_t* hist = nullptr;       // 256 * max_tiles
    uint32_t* dtot = nullptr;       // 256 digit totals
    int64_t capacity = 0;
    int max_tiles = 0;
};

constexpr int RS_THREADS = 256;
constexpr int RS_IPT = 8;
constexpr int RS_TILE = RS_THREADS * RS_IPT;

hipError_t radix_alloc(RadixScratch& s, int64_t n);
void radix_free(RadixScratch& s);
// Sorts the n pairs held in the scratch (s.keys_alt, s.vals_alt; keys <
// 2^bits) into keys/vals, stably.  The error word and the run word are
// checked by every kernel: a nonzero error word or a zero run word (when
// given) turns the sort into a no-op.
hipError_t radix_sort_pairs(RadixScratch& s, uint32_t* keys, uint32_t* vals, int64_t n, int bits,
                            const uint32_t* err, const int32_t* run, hipStream_t stream);

}  // namespace godar
