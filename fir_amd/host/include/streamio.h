// streamio.h -- large binary reads and writes of the host layer (sinogram
// and volume data files) in chunks served by several threads at once, with
// a hook per chunk so that device copies overlap the file transfers. No
// counterpart in the reference, which reads and writes each segment with
// one sequential stream call (ProjInterfileReader.cc:49-94,
// VolInterfileReader.cc:211-273).
#pragma once

#include <cstddef>
#include <functional>
#include <string>

namespace streamio
{
// Called once per chunk (offset and size in bytes within the transfer),
// never concurrently with itself.
using ChunkHook = std::function<void(size_t offset, size_t bytes)>;

// Reads `bytes` bytes from byte `offset` of `path` into `dst`; `landed`
// runs after each chunk has been read. Returns the number of bytes read
// (less than `bytes` only if the file is shorter).
size_t readFile(const std::string& path, size_t offset, void* dst,
                size_t bytes, const ChunkHook& landed = {});

// Writes `bytes` bytes of `src` to `path` (created or truncated); `ready`
// runs before each chunk is written, in chunk order, and may wait until the
// chunk's data are in place.
void writeFile(const std::string& path, const void* src, size_t bytes,
               const ChunkHook& ready = {});
}
