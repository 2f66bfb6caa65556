// szsnn_constants.h -- format constants shared by the device code
// (szsnn_internal.h) and the host-side layout builders (szsnn_layout.h).
#pragma once

#include <stdint.h>

namespace szsnn {

// Pull layout: targets in blocks of 64 * slices (slices from kPullMinSlices
// to kPullSlices, chosen per network), each (block, delay) segment cut into
// slices of 64 targets (one target per lane).  A slice lane's 16-bit word
// holds its target within the block in the low kPullShift bits and, above
// them, how many of its column's 16-byte vectors hold its entries, or
// kPullLaneAll: all of the slice's (the lane's own and the stored padding
// below them, which adds nothing).
constexpr uint32_t kPullShift = 9;
constexpr uint32_t kPullLaneAll = (1u << (16 - kPullShift)) - 1;
constexpr uint32_t kPullTB = 1u << kPullShift;   // targets per pull block, at most
constexpr uint32_t kPullTMask = kPullTB - 1;     // a slice lane's target field
constexpr uint32_t kPullMaxWords = 6144;         // 48 KiB of staged bitmap (sources spanning 393k ids)
constexpr uint32_t kPullSlices = kPullTB / 64;   // slices of 64 targets per (block, delay) segment, at most
constexpr uint32_t kPullMinSlices = 4;           // at least (no fewer than a 256-thread block's waves)
constexpr uint32_t kPullMaxWeights = 255;        // distinct nonzero weights of narrow entries
constexpr uint32_t kMaxPullGroups = 8;           // delay groups of the pull (work items per target block)

// Pull entry formats.  Every entry locates its source in the block's staged
// spike bitmap by rel = source - 64 * (first staged word): bit rel % 32 of
// 32-bit word rel / 32.
//   wide:  8 bytes (rel, fixed-point weight);
//   table: 4 bytes rel | i << kPullRelBits, weight i (1-based) of a table of
//          at most kPullMaxWeights;
//   count: 2 bytes tag | word delta << 2 | bit << 11, for networks with at
//          most kPullCountFields distinct weights.  A lane's entries come in
//          ascending source order; the 32-bit word of an entry's source is
//          the running sum of the word deltas (from word 0 at the start of
//          each column) and `bit` its bit in that word (rel % 32).  Tag
//          i >= 1 is weight i of the table: the lane counts selected entries
//          in kPullCountBits-bit fields of one register, tag i's field at
//          shift kPullCountBits * i.  Tag 0 (the table's inert weight 0) is
//          a skip: a word gap wider than kPullMaxDelta, or padding, which is
//          all zero bits (delta 0), so a zero vector is inert; its field,
//          the lowest, is never read.
// Wide and table padding entries point at a zero word staged after the
// bitmap (rel = 64 * staged words), weight 0 or index 0; an all-zero wide or
// table entry (word 0, weight 0 or index 0) is inert too.
enum : uint32_t { kPullWide = 0, kPullCount = 1, kPullTable = 2 };
constexpr uint32_t kPullRelBits = 21;
constexpr uint32_t kPullRelMask = (1u << kPullRelBits) - 1;
constexpr uint32_t kPullCountFields = 3;
constexpr uint32_t kPullCountBits = 8;
constexpr uint32_t kPullDeltaBits = 9;                       // word delta of a count entry
constexpr uint32_t kPullMaxDelta = (1u << kPullDeltaBits) - 1;
constexpr uint16_t kPullSkipTag = 0;
// Count-format entries are read through a buffer resource with 32-bit byte
// offsets; a lane past its own vectors reads at offset kPullOutOfRange and
// gets a zero vector (no memory access), so the array stays below it.
constexpr uint32_t kPullOutOfRange = 0x80000000u;
constexpr uint64_t kPullMaxBufferBytes = 0x7ff00000ull;
// A count-format entry and its fields.
constexpr uint16_t count_entry(uint32_t tag, uint32_t bit, uint32_t word_delta) {
  return (uint16_t)(tag | word_delta << 2 | bit << 11);
}
constexpr uint32_t count_tag(uint16_t x) { return x & 3u; }
constexpr uint32_t count_delta(uint16_t x) { return (x >> 2) & kPullMaxDelta; }
constexpr uint32_t count_bit(uint16_t x) { return (uint32_t)x >> 11; }
static_assert((uint64_t)(kPullMaxWords + 1) * 64 <= kPullRelMask, "rel of the zero word must fit");
static_assert(kPullMaxWeights < (1u << (32 - kPullRelBits)), "weight index must fit");
static_assert(kPullCountBits * (kPullCountFields + 1) <= 32, "the skip field and the count fields fit one register");
static_assert(kPullSkipTag == 0 && count_entry(kPullSkipTag, 0, 0) == 0, "a zero entry is a skip by 0");
static_assert(kPullCountFields <= 3, "tags are 2 bits");
static_assert(kPullCountBits == 8, "the kernel forms a tag's field shift as tag * 8");
static_assert(2 + 5 + kPullDeltaBits == 16, "count entries are 16 bits");

// Push entries carry the API delay in the target word's top bits.
constexpr uint32_t kSynDelayShift = 26;
constexpr uint32_t kSynTargetMask = (1u << kSynDelayShift) - 1;
constexpr uint32_t kMaxNeurons = 1u << kSynDelayShift;

}  // namespace szsnn
