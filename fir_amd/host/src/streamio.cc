#include <console.h>
#include <streamio.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstring>
#include <fcntl.h>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <vector>

namespace streamio
{
namespace
{
constexpr size_t kChunk = size_t(8) << 20; // bytes per chunk
constexpr unsigned kMaxThreads = 8;

unsigned threadsFor(size_t bytes)
{
  const size_t chunks = (bytes + kChunk - 1) / kChunk;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return static_cast<unsigned>(
    std::min<size_t>({chunks, size_t(kMaxThreads), size_t(hw)}));
}

// Runs body(k) for chunks k = 0 .. n-1 on `threads` threads, each taking
// the next chunk in order; the first exception is rethrown.
template <class Body>
void forChunks(size_t n, unsigned threads, Body body)
{
  std::atomic<size_t> next{0};
  std::exception_ptr failure;
  std::mutex fm;
  auto work = [&] {
    for (size_t k; (k = next.fetch_add(1)) < n;)
    {
      try
      {
        body(k);
      }
      catch (...)
      {
        std::lock_guard<std::mutex> lock(fm);
        if (!failure) failure = std::current_exception();
        next = n;
      }
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  if (failure) std::rethrow_exception(failure);
}
}

size_t readFile(const std::string& path, size_t offset, void* dst,
                size_t bytes, const ChunkHook& landed)
{
  const int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) error("Couldn't open file ", path);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  std::mutex hook;
  std::atomic<size_t> total{0};
  try
  {
    forChunks(n, threadsFor(bytes), [&](size_t k) {
      const size_t off = k * kChunk;
      const size_t len = std::min(kChunk, bytes - off);
      size_t got = 0;
      while (got < len)
      {
        const ssize_t r =
          ::pread(fd, static_cast<char*>(dst) + off + got, len - got,
                  static_cast<off_t>(offset + off + got));
        if (r < 0 && errno == EINTR) continue;
        if (r < 0) error("Couldn't read file ", path, ": ", std::strerror(errno));
        if (r == 0) break;
        got += static_cast<size_t>(r);
      }
      total += got;
      if (landed && got > 0)
      {
        std::lock_guard<std::mutex> lock(hook);
        landed(off, got);
      }
    });
  }
  catch (...)
  {
    ::close(fd);
    throw;
  }
  ::close(fd);
  return total;
}

void writeFile(const std::string& path, const void* src, size_t bytes,
               const ChunkHook& ready)
{
  const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) error("Couldn't create file ", path);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  // chunks are handed out in order, and `ready` runs in that order too
  std::mutex hook;
  size_t next_ready = 0;
  try
  {
    forChunks(n, threadsFor(bytes), [&](size_t k) {
      const size_t off = k * kChunk;
      const size_t len = std::min(kChunk, bytes - off);
      if (ready)
      {
        std::unique_lock<std::mutex> lock(hook);
        while (next_ready <= k)
        {
          const size_t o = next_ready * kChunk;
          ready(o, std::min(kChunk, bytes - o));
          ++next_ready;
        }
      }
      size_t put = 0;
      while (put < len)
      {
        const ssize_t w =
          ::pwrite(fd, static_cast<const char*>(src) + off + put, len - put,
                   static_cast<off_t>(off + put));
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0)
          error("Couldn't write file ", path, ": ", std::strerror(errno));
        put += static_cast<size_t>(w);
      }
    });
  }
  catch (...)
  {
    ::close(fd);
    throw;
  }
  if (::close(fd) != 0) error("Couldn't write file ", path);
}
}
