#include <console.h>
#include <device.h>
#include <reconAlgos.h>

#include <cstdlib>
#include <cstring>
#include <string>

namespace reconAlgos
{
namespace
{
// FIR_PATH_CACHE = snorm16 | fp16 | fp32 selects the path-cache projector
// (Siddon paths kept in device memory) with that length format; unset, empty
// or "off" keeps the on-the-fly walk.
int pathCacheFormat()
{
  const char* env = std::getenv("FIR_PATH_CACHE");
  if (!env || !*env || std::strcmp(env, "off") == 0) return -1;
  if (std::strcmp(env, "snorm16") == 0) return FIR_LEN_SNORM16;
  if (std::strcmp(env, "fp16") == 0) return FIR_LEN_FP16;
  if (std::strcmp(env, "fp32") == 0) return FIR_LEN_FP32;
  error("FIR_PATH_CACHE must be snorm16, fp16, fp32 or off (got \"", env,
        "\")");
  return -1;
}

// FIR_PATH_CACHE_SUBSETS = n caches subsets 0 .. n-1 only and
// FIR_PATH_CACHE_GB = g caps the cache's device memory at g GB; either
// makes the cache partial (fir_cache_create_partial), and so does a whole
// cache that does not fit. Subsets outside the cache use the on-the-fly walk.
long long envNumber(const char* name, double scale)
{
  const char* env = std::getenv(name);
  if (!env || !*env) return 0;
  char* end = nullptr;
  const double d = std::strtod(env, &end);
  if (end == env || *end || d <= 0)
    error(name, " must be a positive number (got \"", env, "\")");
  return static_cast<long long>(d * scale);
}

} // namespace

PathCacheHandle::~PathCacheHandle()
{
  if (cache) fir_cache_destroy(cache);
}

bool pathCacheRequested() { return pathCacheFormat() >= 0; }

std::unique_ptr<PathCacheHandle> buildPathCache(const fir_vol_desc& v,
                                                int nSubsets)
{
  const int cacheFormat = pathCacheFormat();
  if (cacheFormat < 0) return nullptr;
  fir_ctx* ctx = device::context();
  auto h = std::make_unique<PathCacheHandle>();
  print("Building the path cache");
  PhaseTimer tBuild("path cache build");
  // Cache rows per view: tangential bin, then pair by centre slice.
  device::check(fir_set_lor_order(
    ctx, FIR_ORDER_VIEW_MAJOR | FIR_ORDER_TANG_MAJOR | FIR_ORDER_ZCENTER));
  const long long maxSubsets = envNumber("FIR_PATH_CACHE_SUBSETS", 1.0);
  const long long maxBytes = envNumber("FIR_PATH_CACHE_GB", 1e9);
  int rc = FIR_E_NOMEM;
  if (maxSubsets == 0 && maxBytes == 0)
    rc = fir_cache_create(ctx, &v, nSubsets, cacheFormat, &h->cache);
  if (rc == FIR_E_NOMEM)
    rc = fir_cache_create_partial(ctx, &v, nSubsets, cacheFormat,
                                  static_cast<int>(maxSubsets), maxBytes,
                                  &h->cache);
  device::check(fir_set_lor_order(ctx, FIR_ORDER_SEGMENT));
  device::check(rc);
  device::check(fir_synchronize(ctx));
  tBuild.stop();
  device::check(fir_cache_subsets(h->cache, nullptr, &h->nCached));
  if (h->nCached < nSubsets)
    print("The path cache holds subsets 1 to ", h->nCached, " of ", nSubsets,
          "; the others are walked on the fly");
  return h;
}

namespace
{

// Shared body of OSEM and OSEM_ResoReco on device-resident data. With reso
// (reconAlgos.cc:203-367): every sensitivity frame is convolved once (in
// place, in.sens), and each sub-iteration convolves the backprojection and
// updates x /= sens_s, x *= bp (clamped) instead of bp /= sens_s,
// x *= bp.
void osemCore(const ProjData& inputProj, const ScannerData& scanner,
              VolData& outputVol, const std::string& outputVolFileName,
              const OSEMCoreParams& p, const DeviceOSEMInputs& in, bool reso)
{
  scanner.checkProjData(inputProj);
  inputProj.checkNSubsets(p.nSubsets);
  if (p.fwhmXYZ.size() != 3)
    error("\"convolution FHWM XYZ in mm\" must contain three elements");
  if (!in.y || !in.sens) error("OSEM: device sinogram or sensitivity missing");

  const bool convolveFlag = p.convolutionInterval > 0 && p.fwhmXYZ[0] > 0.0 &&
    p.fwhmXYZ[1] > 0.0 && p.fwhmXYZ[2] > 0.0;

  device::useGeometry(scanner, inputProj.getHeader());
  fir_ctx* ctx = device::context();
  const fir_vol_desc v = device::describe(outputVol.getHeader());
  const size_t nv = static_cast<size_t>(outputVol.getNVoxelsPerFrame());
  const size_t vb = nv * sizeof(float);

  struct DevView
  {
    const float* p;
    const float* f() const { return p; }
  } y{in.y}, bias{in.bias};
  struct SensView
  {
    float* p;
    float* f() const { return p; }
  } s{in.sens};
  const bool hasBias = in.bias != nullptr;
  device::Buffer x(vb), bp(vb);
  x.upload(outputVol.getDataArray(), vb);
  device::check(fir_fill(ctx, bp.f(), static_cast<int64_t>(nv), 0.0f));
  device::check(fir_cut_circle(ctx, &v, x.f(), 1, p.cutRadius));

  // the path cache handed in, or one built here when FIR_PATH_CACHE asks
  std::unique_ptr<PathCacheHandle> own;
  struct
  {
    fir_cache* cache = nullptr;
  } cached;
  int nCached = 0;
  if (in.cache)
  {
    cached.cache = in.cache;
    device::check(fir_cache_subsets(in.cache, nullptr, &nCached));
  }
  else if ((own = buildPathCache(v, p.nSubsets)))
  {
    cached.cache = own->cache;
    nCached = own->nCached;
  }
  if (in.computeSens)
  {
    // the sensitivity of the projector the iterations use: the cached
    // subsets' frames from the cache, the others from the walk
    PhaseTimer tSens("sensitivity");
    echo("Sensitivity (GPU)");
    if (nCached < p.nSubsets)
      device::check(fir_sensitivity(ctx, &v, s.f(), p.nSubsets));
    if (cached.cache)
      device::check(fir_cache_sensitivity(cached.cache, s.f()));
    if (in.sensOut)
      device::check(fir_memcpy_d2h(ctx, in.sensOut->getAllFrames(), s.f(),
                                   vb * p.nSubsets));
    device::check(fir_synchronize(ctx));
  }
  if (reso)
    device::check(fir_convolve(ctx, &v, s.f(), p.nSubsets, p.fwhmXYZ.data(),
                               p.cutRadius));
  auto projector = [&](int subset) {
    if (cached.cache && subset < nCached)
    {
      const float* xs[1] = {x.f()};
      const float* ys[1] = {y.f()};
      const float* bs[1] = {hasBias ? bias.f() : nullptr};
      float* bps[1] = {bp.f()};
      device::check(
        fir_cache_osem_project(cached.cache, 1, xs, ys, bs, bps, subset));
    }
    else
      device::check(fir_osem_project(ctx, &v, x.f(), y.f(),
                                     hasBias ? bias.f() : nullptr, bp.f(),
                                     subset, p.nSubsets));
  };

  const int total = p.nIterations * p.nSubsets;
  for (int iter = 0; iter < p.nIterations; ++iter)
  {
    print("Iteration ", iter + 1, " of ", p.nIterations);
    for (int subset = 0; subset < p.nSubsets; ++subset)
    {
      const int subiter = iter * p.nSubsets + subset + 1;
      if (p.nSubsets > 1)
        print("  Sub-iteration ", subset + 1, " of ", p.nSubsets);
      const int conv =
        convolveFlag && subiter % p.convolutionInterval == 0 ? 1 : 0;
      const int reset = subiter != total ? 1 : 0;
      float* sens_s = s.f() + nv * subset;
      if (reso)
      {
        // The reference also convolves a copy of x here that no later step
        // reads (reconAlgos.cc:246-247, 271-274, 349-353); it is skipped.
        projector(subset);
        device::check(
          fir_convolve(ctx, &v, bp.f(), 1, p.fwhmXYZ.data(), p.cutRadius));
        device::check(fir_div_clamped(ctx, x.f(), sens_s, (int64_t)nv));
        device::check(fir_mul_clamped(ctx, x.f(), bp.f(), (int64_t)nv));
        if (reset) device::check(fir_fill(ctx, bp.f(), (int64_t)nv, 0.0f));
        if (conv)
          device::check(
            fir_convolve(ctx, &v, x.f(), 1, p.fwhmXYZ.data(), p.cutRadius));
        device::check(fir_cut_circle(ctx, &v, x.f(), 1, p.cutRadius));
      }
      else if (cached.cache && subset < nCached)
      {
        projector(subset);
        device::check(fir_osem_update(ctx, &v, x.f(), bp.f(), sens_s,
                                      p.cutRadius, conv ? 0 : 1, reset));
        if (conv)
        {
          device::check(
            fir_convolve(ctx, &v, x.f(), 1, p.fwhmXYZ.data(), p.cutRadius));
          device::check(fir_cut_circle(ctx, &v, x.f(), 1, p.cutRadius));
        }
      }
      else
        device::check(fir_osem_subiter(
          ctx, &v, x.f(), y.f(), hasBias ? bias.f() : nullptr, sens_s,
          bp.f(), subset, p.nSubsets, conv, p.fwhmXYZ.data(), p.cutRadius,
          reset));
      if (p.saveInterval > 0 && subiter % p.saveInterval == 0 &&
          subiter != total)
      {
        x.download(outputVol.getDataArray(), vb);
        outputVol.write(outputVolFileName + "_subiter_" +
                        std::to_string(subiter));
      }
    }
  }
  x.download(outputVol.getDataArray(), vb);
}

// The reference API's entry: host sensitivity and sinograms, uploaded; the
// convolved sensitivity of OSEM_ResoReco is written back to sensOut.
void osemImpl(const ProjData& inputProj, const ScannerData& scanner,
              VolData& outputVol, const std::string& outputVolFileName,
              const OSEMCoreParams& p, const VolData& sens,
              const std::optional<ProjData>& biasProj, bool reso,
              VolData* sensOut)
{
  scanner.checkProjData(inputProj);
  inputProj.checkNSubsets(p.nSubsets);
  if (sens.getNFrames() != p.nSubsets)
    error("Wrong number of frames (expected ", p.nSubsets, ")");
  if (sens.getHeader() != outputVol.getHeader())
    error("Sensitivity volume provided doesn't fit with output volume "
          "provided");
  if (biasProj && biasProj->getHeader() != inputProj.getHeader())
    error("Bias projection doesn't fit with input projection");
  const size_t vb =
    static_cast<size_t>(outputVol.getNVoxelsPerFrame()) * sizeof(float);
  const size_t yb =
    static_cast<size_t>(inputProj.getGeometry().nBins) * sizeof(float);
  device::Buffer y(yb), s(vb * p.nSubsets), bias;
  y.upload(inputProj.getContiguousData(), yb);
  s.upload(sens.getAllFrames(), vb * p.nSubsets);
  if (biasProj)
  {
    bias = device::Buffer(yb);
    bias.upload(biasProj->getContiguousData(), yb);
  }
  DeviceOSEMInputs in;
  in.y = y.f();
  in.bias = biasProj ? bias.f() : nullptr;
  in.sens = s.f();
  osemCore(inputProj, scanner, outputVol, outputVolFileName, p, in, reso);
  if (sensOut) s.download(sensOut->getAllFrames(), vb * p.nSubsets);
}
}

void OSEM(const ProjData& inputProj, const ScannerData& scanner,
          VolData& outputVol, const std::string& outputVolFileName,
          const OSEMCoreParams& p, const VolData& sens,
          const std::optional<ProjData>& biasProj)
{
  echo("OSEM:");
  osemImpl(inputProj, scanner, outputVol, outputVolFileName, p, sens,
           biasProj, false, nullptr);
}

void OSEMOnDevice(const ProjData& inputProj, const ScannerData& scanner,
                  VolData& outputVol, const std::string& outputVolFileName,
                  const OSEMCoreParams& p, const DeviceOSEMInputs& in,
                  bool reso)
{
  echo(reso ? "OSEM_ResoReco:" : "OSEM:");
  osemCore(inputProj, scanner, outputVol, outputVolFileName, p, in, reso);
}

void OSEM_ResoReco(const ProjData& inputProj, const ScannerData& scanner,
                   VolData& outputVol, const std::string& outputVolFileName,
                   const OSEMCoreParams& p, VolData& sensitivityMap,
                   const std::optional<ProjData>& biasProj)
{
  echo("OSEM_ResoReco:");
  osemImpl(inputProj, scanner, outputVol, outputVolFileName, p,
           sensitivityMap, biasProj, true, &sensitivityMap);
}
}
