// pinned.h -- allocator of the host containers' data (sinograms, volumes):
// page-locked memory from fir_host_alloc, so that uploads to and downloads
// from the GPU run as direct DMA. Where page-locked memory cannot be had
// (no device present, or FIR_PINNED=0) it hands out ordinary memory; the
// containers and their transfers work the same either way.
#pragma once

#include <cstddef>
#include <new>

namespace pinned
{
// Page-locked block of at least `bytes`, or ordinary memory; never NULL
// (throws std::bad_alloc).
void* acquire(std::size_t bytes);
void release(void* p);
// Whether p came from page-locked memory.
bool isPinned(const void* p);

template <class T>
struct Allocator
{
  using value_type = T;
  Allocator() = default;
  template <class U>
  Allocator(const Allocator<U>&) noexcept
  {
  }
  T* allocate(std::size_t n) { return static_cast<T*>(acquire(n * sizeof(T))); }
  void deallocate(T* p, std::size_t) noexcept { release(p); }
  template <class U>
  bool operator==(const Allocator<U>&) const noexcept
  {
    return true;
  }
  template <class U>
  bool operator!=(const Allocator<U>&) const noexcept
  {
    return false;
  }
};
}
