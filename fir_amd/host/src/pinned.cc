#include <pinned.h>

#include <fir_hip.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>

namespace pinned
{
namespace
{
// Every block starts with a header that records where it came from; the
// data follow at a 64-byte offset (which keeps them 64-byte aligned).
constexpr std::size_t kHeader = 64;
constexpr std::uint32_t kPinned = 0x50494e4e;   // "PINN"
constexpr std::uint32_t kPageable = 0x50414745; // "PAGE"

bool wanted()
{
  static const bool on = [] {
    const char* e = std::getenv("FIR_PINNED");
    return !(e && std::strcmp(e, "0") == 0);
  }();
  return on;
}
}

void* acquire(std::size_t bytes)
{
  void* base = nullptr;
  std::uint32_t kind = kPinned;
  if (!wanted() || fir_host_alloc(&base, bytes + kHeader) != FIR_OK)
  {
    kind = kPageable;
    base = std::aligned_alloc(kHeader, ((bytes + 2 * kHeader - 1) / kHeader) *
                                         kHeader);
    if (!base) throw std::bad_alloc();
  }
  std::memcpy(base, &kind, sizeof(kind));
  return static_cast<char*>(base) + kHeader;
}

void release(void* p)
{
  if (!p) return;
  char* base = static_cast<char*>(p) - kHeader;
  std::uint32_t kind;
  std::memcpy(&kind, base, sizeof(kind));
  if (kind == kPinned)
    (void)fir_host_free(base);
  else
    std::free(base);
}

bool isPinned(const void* p)
{
  if (!p) return false;
  std::uint32_t kind;
  std::memcpy(&kind, static_cast<const char*>(p) - kHeader, sizeof(kind));
  return kind == kPinned;
}
}
