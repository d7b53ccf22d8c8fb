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

struct CacheHolder
{
  fir_cache* cache = nullptr;
  ~CacheHolder()
  {
    if (cache) fir_cache_destroy(cache);
  }
};
}

void OSEM(const ProjData& inputProj, const ScannerData& scanner,
          VolData& outputVol, const std::string& outputVolFileName,
          const OSEMCoreParams& p, const VolData& sens,
          const std::optional<ProjData>& biasProj)
{
  echo("OSEM:");
  scanner.checkProjData(inputProj);
  inputProj.checkNSubsets(p.nSubsets);
  if (sens.getNFrames() != p.nSubsets)
    error("Wrong number of frames (expected ", p.nSubsets, ")");
  if (sens.getHeader() != outputVol.getHeader())
    error("Sensitivity volume provided doesn't fit with output volume "
          "provided");
  if (biasProj && biasProj->getHeader() != inputProj.getHeader())
    error("Bias projection doesn't fit with input projection");
  if (p.fwhmXYZ.size() != 3)
    error("\"convolution FHWM XYZ in mm\" must contain three elements");

  const bool convolveFlag = p.convolutionInterval > 0 && p.fwhmXYZ[0] > 0.0 &&
    p.fwhmXYZ[1] > 0.0 && p.fwhmXYZ[2] > 0.0;

  device::useGeometry(scanner, inputProj.getHeader());
  fir_ctx* ctx = device::context();
  const fir_vol_desc v = device::describe(outputVol.getHeader());
  const size_t nv = static_cast<size_t>(outputVol.getNVoxelsPerFrame());
  const size_t nb = static_cast<size_t>(inputProj.getGeometry().nBins);
  const size_t vb = nv * sizeof(float);

  device::Buffer x(vb), bp(vb), y(nb * sizeof(float));
  device::Buffer s(vb * p.nSubsets), bias;
  x.upload(outputVol.getDataArray(), vb);
  y.upload(inputProj.getContiguousData(), nb * sizeof(float));
  s.upload(sens.getAllFrames(), vb * p.nSubsets);
  if (biasProj)
  {
    bias = device::Buffer(nb * sizeof(float));
    bias.upload(biasProj->getContiguousData(), nb * sizeof(float));
  }
  device::check(fir_fill(ctx, bp.f(), static_cast<int64_t>(nv), 0.0f));
  device::check(fir_cut_circle(ctx, &v, x.f(), 1, p.cutRadius));

  CacheHolder cached;
  const int cacheFormat = pathCacheFormat();
  if (cacheFormat >= 0)
  {
    print("Building the path cache");
    // Cache rows per view: tangential bin, then pair by centre slice.
    device::check(fir_set_lor_order(
      ctx, FIR_ORDER_VIEW_MAJOR | FIR_ORDER_TANG_MAJOR | FIR_ORDER_ZCENTER));
    const int rc =
      fir_cache_create(ctx, &v, p.nSubsets, cacheFormat, &cached.cache);
    device::check(fir_set_lor_order(ctx, FIR_ORDER_SEGMENT));
    device::check(rc);
  }

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
      if (cached.cache)
      {
        const float* xs[1] = {x.f()};
        const float* ys[1] = {y.f()};
        const float* bs[1] = {biasProj ? bias.f() : nullptr};
        float* bps[1] = {bp.f()};
        device::check(
          fir_cache_osem_project(cached.cache, 1, xs, ys, bs, bps, subset));
        device::check(fir_osem_update(ctx, &v, x.f(), bp.f(),
                                      s.f() + nv * subset, p.cutRadius,
                                      conv ? 0 : 1, subiter != total ? 1 : 0));
        if (conv)
        {
          device::check(
            fir_convolve(ctx, &v, x.f(), 1, p.fwhmXYZ.data(), p.cutRadius));
          device::check(fir_cut_circle(ctx, &v, x.f(), 1, p.cutRadius));
        }
      }
      else
        device::check(fir_osem_subiter(
          ctx, &v, x.f(), y.f(), biasProj ? bias.f() : nullptr,
          s.f() + nv * subset, bp.f(), subset, p.nSubsets, conv,
          p.fwhmXYZ.data(), p.cutRadius, subiter != total ? 1 : 0));
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
}
