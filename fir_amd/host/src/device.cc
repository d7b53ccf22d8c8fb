#include <ProjHeader.h>
#include <ScannerData.h>
#include <console.h>
#include <device.h>
#include <ProjData.h>
#include <ProjInterfileReader.h>

#include <cstdlib>
#include <cstring>
#include <vector>

namespace device
{
namespace
{
struct State
{
  fir_ctx* ctx = nullptr;
  // geometry currently loaded in ctx
  std::vector<double> xy, z;
  ProjHeader proj{};
  bool projSet = false;

  ~State()
  {
    if (ctx) fir_ctx_destroy(ctx);
  }
};

State& state()
{
  static State s;
  return s;
}
}

void check(int rc)
{
  if (rc != FIR_OK) error(fir_last_error());
}

fir_ctx* context()
{
  State& s = state();
  if (!s.ctx)
  {
    PhaseTimer t("device context");
    const char* env = std::getenv("FIR_DEVICE");
    const int dev = env ? std::atoi(env) : 0;
    check(fir_ctx_create(&s.ctx, dev));
  }
  return s.ctx;
}

Buffer::Buffer(size_t bytes) : mBytes(bytes)
{
  if (bytes) check(fir_malloc(context(), &mPtr, bytes));
}

Buffer::~Buffer()
{
  if (mPtr) fir_free(context(), mPtr);
}

Buffer::Buffer(Buffer&& o) noexcept : mPtr(o.mPtr), mBytes(o.mBytes)
{
  o.mPtr = nullptr;
  o.mBytes = 0;
}

Buffer& Buffer::operator=(Buffer&& o) noexcept
{
  if (this != &o)
  {
    if (mPtr) fir_free(context(), mPtr);
    mPtr = o.mPtr;
    mBytes = o.mBytes;
    o.mPtr = nullptr;
    o.mBytes = 0;
  }
  return *this;
}

void Buffer::upload(const void* host, size_t bytes)
{
  if (bytes > mBytes) error("device buffer too small for upload");
  if (bytes) check(fir_memcpy_h2d(context(), mPtr, host, bytes));
}

void Buffer::download(void* host, size_t bytes) const
{
  if (bytes > mBytes) error("device buffer too small for download");
  if (bytes) check(fir_memcpy_d2h(context(), host, mPtr, bytes));
}

Buffer readProj(const std::string& headerFile, ProjData& proj)
{
  proj.read(headerFile, ProjData::ConstructionMode::ALLOCATE);
  const size_t bytes =
    static_cast<size_t>(proj.getGeometry().nBins) * sizeof(types::BinValue);
  Buffer dev(bytes);
  fir_ctx* c = context();
  char* host = reinterpret_cast<char*>(proj.getContiguousData());
  char* d = static_cast<char*>(dev.get());
  ProjInterfileReader reader(headerFile);
  reader.readData(proj.getDataArray(), [&](size_t off, size_t n) {
    check(fir_memcpy_h2d_async(c, d + off, host + off, n));
  });
  check(fir_synchronize(c));
  return dev;
}

void writeProj(const Buffer& dev, ProjData& proj,
               const std::string& outputProjFile)
{
  const size_t bytes =
    static_cast<size_t>(proj.getGeometry().nBins) * sizeof(types::BinValue);
  if (dev.size() < bytes) error("device sinogram too small for the header");
  fir_ctx* c = context();
  char* host = reinterpret_cast<char*>(proj.getContiguousData());
  const char* d = static_cast<const char*>(dev.get());
  ProjInterfileReader::writeProjInterfile(
    outputProjFile, proj.getHeader(), proj.getDataArray(),
    [&](size_t off, size_t n) {
      check(fir_memcpy_d2h_async(c, host + off, d + off, n));
      check(fir_synchronize(c));
    });
}

fir_vol_desc describe(const VolHeader& h)
{
  fir_vol_desc d;
  d.nx = h.volSize.nPixelsX;
  d.ny = h.volSize.nPixelsY;
  d.nz = h.volSize.nSlices;
  d.dx = h.voxelExtent.pixelWidth;
  d.dy = h.voxelExtent.pixelHeight;
  d.dz = h.voxelExtent.sliceThickness;
  d.ox = h.volOffset.x;
  d.oy = h.volOffset.y;
  d.oz = h.volOffset.z;
  return d;
}

void useGeometry(const ScannerData& scanner, const ProjHeader& proj)
{
  State& s = state();
  fir_ctx* c = context();
  const int n = scanner.getGeometry().nCrystalsPerRing;
  const int ns = scanner.getGeometry().nSlices;
  std::vector<double> xy(2 * static_cast<size_t>(n));
  const auto* p = scanner.getCrystalXYPositionVector();
  for (int i = 0; i < n; ++i)
  {
    xy[2 * i] = p[i].x;
    xy[2 * i + 1] = p[i].y;
  }
  std::vector<double> z(scanner.getSliceZPositionVector(),
                        scanner.getSliceZPositionVector() + ns);
  if (xy != s.xy || z != s.z)
  {
    check(fir_set_scanner(c, xy.data(), n, z.data(), ns));
    s.xy = xy;
    s.z = z;
  }
  useProj(proj);
}

void useProj(const ProjHeader& proj)
{
  State& s = state();
  fir_ctx* c = context();
  if (!s.projSet || s.proj != proj)
  {
    check(fir_set_proj(c, proj.nRings, proj.nCrystalsPerRing,
                       proj.segmentSpan, proj.nSegments, proj.nTangCoords));
    s.proj = proj;
    s.projSet = true;
  }
}

void clampedOp(ClampedOp op, float* a, const float* b, long long n)
{
  if (n <= 0) return;
  const size_t bytes = sizeof(float) * static_cast<size_t>(n);
  Buffer da(bytes), db(bytes);
  da.upload(a, bytes);
  db.upload(b, bytes);
  check(op == ClampedOp::Mul ? fir_mul_clamped(context(), da.f(), db.f(), n)
                             : fir_div_clamped(context(), da.f(), db.f(), n));
  da.download(a, bytes);
}

void expClamped(float* a, long long n)
{
  if (n <= 0) return;
  const size_t bytes = sizeof(float) * static_cast<size_t>(n);
  Buffer da(bytes);
  da.upload(a, bytes);
  check(fir_exp_clamped(context(), da.f(), n));
  da.download(a, bytes);
}

void rebinWeight(const ProjHeader& proj, float* a, long long n)
{
  if (n <= 0 || proj.segmentSpan == 1) return;
  useProj(proj);
  const size_t bytes = sizeof(float) * static_cast<size_t>(n);
  Buffer da(bytes);
  da.upload(a, bytes);
  check(fir_rebin_weight(context(), da.f()));
  da.download(a, bytes);
}

void huToMu(float* a, long long n)
{
  if (n <= 0) return;
  const size_t bytes = sizeof(float) * static_cast<size_t>(n);
  Buffer da(bytes);
  da.upload(a, bytes);
  check(fir_hu_to_mu(context(), da.f(), n));
  da.download(a, bytes);
}
}
