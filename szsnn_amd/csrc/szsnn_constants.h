// szsnn_constants.h -- format constants shared by the device code
// (szsnn_internal.h) and the host-side layout builders (szsnn_layout.h).
#pragma once

#include <stdint.h>

namespace szsnn {

// Pull layout: targets in blocks of kPullTB, each (block, delay) segment cut
// into slices of 64 targets (one target per lane).
constexpr uint32_t kPullShift = 9;
constexpr uint32_t kPullTB = 1u << kPullShift;   // targets per pull block
constexpr uint32_t kPullMaxWords = 6144;         // 48 KiB of staged bitmap (sources spanning 393k ids)
constexpr uint32_t kPullSlices = kPullTB / 64;   // slices of 64 targets per (block, delay) segment
constexpr uint32_t kPullMaxWeights = 255;        // distinct nonzero weights of narrow entries
constexpr uint32_t kMaxPullGroups = 8;           // delay groups of the pull (work items per target block)

// Push entries carry the API delay in the target word's top bits.
constexpr uint32_t kSynDelayShift = 26;
constexpr uint32_t kSynTargetMask = (1u << kSynDelayShift) - 1;
constexpr uint32_t kMaxNeurons = 1u << kSynDelayShift;

}  // namespace szsnn
