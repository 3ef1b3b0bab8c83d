// First-fit allocator over one device arena (reference interface and
// behaviour: stream/gear/include/MemoryManager.cuh:8-170): segments are kept
// in address order, requests are rounded up to 256 B (a zero request takes
// 8 B), freeing coalesces with free neighbours, exhaustion throws
// StrException, and the high-water mark is tracked.
#pragma once

#include <algorithm>
#include <cstdio>
#include <list>
#include <string>

#include "Common.h"

struct MemoryManager {
  static constexpr size_t guarantee_alignment = 256;

  struct MemorySegment {
    size_t start;
    size_t size;
    std::string tag; // "" = free
  };

  size_t max_available_memory = 0;
  std::list<MemorySegment> memory_segments {{0, 0, ""}};
  size_t total_memory_required = 0;

  void set_reserved_memory(size_t bytes)
  {
    max_available_memory = bytes;
    free_all();
  }

  static size_t aligned(size_t requested)
  {
    if (requested == 0) requested = 8;
    return (requested + guarantee_alignment - 1) / guarantee_alignment * guarantee_alignment;
  }

  // Returns the byte offset of a new segment tagged `tag`.
  size_t reserve(const std::string& tag, size_t requested)
  {
    const size_t request = aligned(requested);
    auto it = std::find_if(memory_segments.begin(), memory_segments.end(), [&](const MemorySegment& s) {
      return s.tag.empty() && s.size >= request;
    });
    if (it == memory_segments.end()) {
      throw StrException(
        "Reserve: Requested size for argument " + tag + " could not be met (" +
        std::to_string((double) request / (1024.0 * 1024.0)) + " MiB)");
    }
    const size_t start = it->start;
    it->start += request;
    it->size -= request;
    if (it->size == 0) it = memory_segments.erase(it);
    memory_segments.insert(it, MemorySegment {start, request, tag});
    const auto& last = memory_segments.back();
    const size_t used_up_to = last.tag.empty() ? last.start : last.start + last.size;
    total_memory_required = std::max(total_memory_required, used_up_to);
    return start;
  }

  void free(const std::string& tag)
  {
    auto it = std::find_if(
      memory_segments.begin(), memory_segments.end(), [&](const MemorySegment& s) { return s.tag == tag; });
    if (it == memory_segments.end()) throw StrException("Free: Requested tag could not be found (" + tag + ")");
    it->tag.clear();
    if (it != memory_segments.begin()) {
      auto prev = std::prev(it);
      if (prev->tag.empty()) {
        prev->size += it->size;
        memory_segments.erase(it);
        it = prev;
      }
    }
    auto next = std::next(it);
    if (next != memory_segments.end() && next->tag.empty()) {
      it->size += next->size;
      memory_segments.erase(next);
    }
  }

  void free_all() { memory_segments = std::list<MemorySegment> {{0, max_available_memory, ""}}; }

  bool is_reserved(const std::string& tag) const
  {
    return std::any_of(
      memory_segments.begin(), memory_segments.end(), [&](const MemorySegment& s) { return s.tag == tag; });
  }

  void print() const
  {
    std::printf("Memory segments (MiB):\n");
    for (const auto& s : memory_segments)
      std::printf("%s (%.3f), ", s.tag.empty() ? "unused" : s.tag.c_str(), (double) s.size / (1024.0 * 1024.0));
    std::printf("\nMax memory required: %.3f MiB\n\n", (double) total_memory_required / (1024.0 * 1024.0));
  }
};
