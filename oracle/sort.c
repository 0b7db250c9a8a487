/*
 * sort.c - CPU restatement of the reference radix sort. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows crates/dynamis-sort/src/radix.rs and its shaders:
 *   units_for     radix.rs:63-67
 *   key_words     radix.rs:71-74
 *   digit order   radix.rs:312-317 (`digit_passes`: minor digits 0.., then major digits 4..)
 *   length        sort_length.wgsl:7 (len = min(count, lane length))
 *   per digit     sort_histogram.wgsl:11-36 (count digits), sort_prefix.wgsl:9-27 (exclusive
 *                 prefix), sort_scatter.wgsl:45-104 (stable scatter in input order)
 *   odd passes    sort_copy.wgsl:12-31 (scratch back to the primaries, radix.rs:472)
 * The reference splits each digit pass over units and tiles; the scatter ranks keys in input
 * order within and across units, which is the serial stable counting sort restated here.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

uint32_t oracle_units_for(uint32_t live) {
    uint64_t units = ((uint64_t)live + 255u) / 256u + 64u; /* TILE=256, PLAN_SLACK=64 */
    if (units < 1u) units = 1u;
    if (units > 512u) units = 512u; /* UNITS */
    return (uint32_t)units;
}

uint32_t oracle_key_words(uint32_t elements) {
    uint32_t last = elements ? elements - 1u : 0u; /* saturating_sub(1) */
    uint32_t bits = 0;
    while (bits < 32u && (last >> bits) != 0u) bits++; /* 32 - leading_zeros */
    uint32_t words = (bits + 7u) / 8u;
    if (words < 1u) words = 1u;
    if (words > 4u) words = 4u;
    return words;
}

int oracle_sort_pairs(uint32_t* major, uint32_t* minor, uint32_t* payload, uint32_t capacity,
                      uint32_t count, uint32_t base, uint32_t major_words, uint32_t minor_words) {
    uint32_t digits[8];
    uint32_t passes = 0;
    for (uint32_t d = 0; d < minor_words && d < 4u; d++) digits[passes++] = d;
    for (uint32_t d = 0; d < major_words && d < 4u; d++) digits[passes++] = 4u + d;
    if (passes == 0) return -1;
    if (base > capacity) base = capacity;
    uint32_t len = count < capacity - base ? count : capacity - base;
    if (len == 0) return 0;

    uint32_t* lanes[2][3];
    lanes[0][0] = major + base;
    lanes[0][1] = minor + base;
    lanes[0][2] = payload + base;
    for (int l = 0; l < 3; l++) lanes[1][l] = (uint32_t*)malloc((size_t)len * 4u);

    for (uint32_t p = 0; p < passes; p++) {
        uint32_t** src = lanes[p % 2];
        uint32_t** dst = lanes[(p + 1) % 2];
        uint32_t digit = digits[p];
        const uint32_t* word = digit >= 4u ? src[0] : src[1];
        uint32_t shift = (digit % 4u) * 8u;
        uint32_t counts[256];
        memset(counts, 0, sizeof counts);
        for (uint32_t i = 0; i < len; i++) counts[(word[i] >> shift) & 255u]++;
        uint32_t offsets[256];
        uint32_t running = 0;
        for (uint32_t b = 0; b < 256u; b++) {
            offsets[b] = running;
            running += counts[b];
        }
        for (uint32_t i = 0; i < len; i++) {
            uint32_t place = offsets[(word[i] >> shift) & 255u]++;
            dst[0][place] = src[0][i];
            dst[1][place] = src[1][i];
            dst[2][place] = src[2][i];
        }
    }
    if (passes % 2u == 1u) {
        for (int l = 0; l < 3; l++) memcpy(lanes[0][l], lanes[1][l], (size_t)len * 4u);
    }
    for (int l = 0; l < 3; l++) free(lanes[1][l]);
    return 0;
}
