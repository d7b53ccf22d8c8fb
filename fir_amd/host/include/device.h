// device.h -- the host layer's link to the GPU: one process-wide fir_ctx
// (HIP device from $FIR_DEVICE, default 0), RAII device buffers, and the
// conversions from host containers to the C-ABI's descriptors.
#pragma once

#include <VolHeader.h>

#include <cstddef>
#include <fir_hip.h>
#include <string>


class ProjData;
class ScannerData;
struct ProjHeader;

namespace device
{
// The context; created on first use. Throws (error()) if no device.
fir_ctx* context();

// Throws error(<fir_last_error()>) if rc != 0.
void check(int rc);

class Buffer
{
public:
  Buffer() = default;
  explicit Buffer(size_t bytes);
  ~Buffer();
  Buffer(Buffer&& o) noexcept;
  Buffer& operator=(Buffer&& o) noexcept;
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;

  void* get() const { return mPtr; }
  float* f() const { return static_cast<float*>(mPtr); }
  size_t size() const { return mBytes; }
  void upload(const void* host, size_t bytes);
  void download(void* host, size_t bytes) const;

private:
  void* mPtr = nullptr;
  size_t mBytes = 0;
};

fir_vol_desc describe(const VolHeader& header);

// Reads the sinogram of `headerFile` into `proj` (as ProjData::read with
// READ_DATA) and into a new device buffer: each chunk's copy to the device
// is enqueued as soon as the chunk has been read, so that the copies run
// while the rest of the file is read (proj's storage is page-locked).
Buffer readProj(const std::string& headerFile, ProjData& proj);

// Writes `proj`'s header and the device sinogram `dev` to outputProjFile
// (as ProjData::write): each chunk is downloaded into proj's storage just
// before it is written, while earlier chunks are still being written.
void writeProj(const Buffer& dev, ProjData& proj,
               const std::string& outputProjFile);

// Make the context's scanner tables and sinogram geometry those of
// (scanner, proj); no-op if they already are.
void useGeometry(const ScannerData& scanner, const ProjHeader& proj);
// The same for the sinogram geometry alone.
void useProj(const ProjHeader& proj);

enum class ClampedOp
{
  Mul,
  Div
};

// Element-wise operations on host arrays, computed on the GPU.
void clampedOp(ClampedOp op, float* a, const float* b, long long n);
void expClamped(float* a, long long n);
void huToMu(float* a, long long n);
// ProjData::rebinWeight of a whole sinogram (n bins of geometry proj).
void rebinWeight(const ProjHeader& proj, float* a, long long n);
}
