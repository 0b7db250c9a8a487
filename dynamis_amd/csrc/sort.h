// Internal interface of the onesweep LSD radix sort (sort.hip), shared with the stages that feed
// it a precomputed digit histogram (the broadphase fuses the histogram into entry emission).
#pragma once

#include "common.h"

namespace dyn {

constexpr int kSortThreads = 256;                      // 4 waves
constexpr int kSortItems = 16;                         // keys per lane per tile
constexpr int kSortTile = kSortThreads * kSortItems;   // 4096 keys per tile
constexpr int kSortMinTile = kSortThreads * 8;        // narrowest tiling (shorter streams)
constexpr int kSortMaxPasses = 8;                      // 4 minor + 4 major digits
constexpr int kSortBins = 256;
constexpr uint32_t kSortMaxCapacity = (1u << 30) - 1;  // look-back values carry 30 bits

// Control words kept in scratch.
enum SortCtl : uint32_t {
    kCtlLen = 0,    // records to sort (read on device)
    kCtlBase = 1,   // first record
    kCtlError = 2,  // non-zero when a look-back gave up waiting (never expected)
    kCtlTiles = 8,  // 8 tile-acquisition counters, one per pass
    kCtlWords = 16,
};

// Digit of one pass: which key lane it reads and the byte shift within that word.
struct SortDigit {
    uint32_t major;  // 1: major lane, 0: minor lane
    uint32_t shift;  // 0, 8, 16 or 24
};

// Kernel-timing names of one sort call site.
struct SortNames {
    const char* prepare;
    const char* histogram;
    const char* pass;
    const char* copy;
};
extern const SortNames kGenericSortNames;  // dyn_sort_pairs
extern const SortNames kEntrySortNames;    // the broadphase's entry-region sort
extern const SortNames kPairSortNames;     // the canonical pair sort

struct SortScratch {
    uint32_t* major = nullptr;    // scratch lanes, `capacity` words each
    uint32_t* minor = nullptr;
    uint32_t* payload = nullptr;
    uint32_t* hist = nullptr;     // kSortMaxPasses * 256 digit counts
    uint32_t* ctl = nullptr;      // kCtlWords control words
    uint32_t* flags = nullptr;    // passes * max_tiles * 256 look-back words
    uint32_t max_tiles = 0;       // tiles of the narrowest tiling
};

// Lays the sort's scratch out over `base` (may be null to only size it); returns the bytes used.
size_t sort_scratch_layout(uint32_t capacity, uint32_t passes, bool minor_lane, void* base,
                           SortScratch* out);

// The digit list of the reference's `digit_passes` (radix.rs:312-317): minor bytes first, then
// major bytes.
int sort_digits(uint32_t major_words, uint32_t minor_words, SortDigit* digits);

// Sets len = min(*d_count, capacity - base), base = *d_base (or 0), and clears the histogram and
// the tile counters. One workgroup.
int sort_prepare(hipStream_t stream, const SortScratch& s, const uint32_t* d_count,
                 const uint32_t* d_base, uint32_t capacity,
                 const SortNames& names = kGenericSortNames);

// Counts every pass's digits in one read of the keys and clears the first pass's look-back words.
int sort_histogram(hipStream_t stream, const uint32_t* major, const uint32_t* minor,
                   const SortScratch& s, const SortDigit* digits, uint32_t passes,
                   uint32_t capacity, const SortNames& names = kGenericSortNames);

// Runs the digit passes (histogram already in s.hist, first pass's look-back words clear) and
// leaves the result in the primaries. `minor` may be null when the minor lane is not carried.
int sort_passes(hipStream_t stream, uint32_t* major, uint32_t* minor, uint32_t* payload,
                const SortScratch& s, const SortDigit* digits, uint32_t passes,
                uint32_t capacity, const SortNames& names = kGenericSortNames);

}  // namespace dyn
