#include <console.h>
#include <device.h>
#include <operations.h>

namespace operations
{
void convolve(VolData& vol, std::vector<float> fwhm, float cutRadius)
{
  if (fwhm.size() != 3) error("convolution FWHM must have three elements");
  if (!(fwhm[0] > 0.0 && fwhm[1] > 0.0 && fwhm[2] > 0.0)) return;
  const fir_vol_desc v = device::describe(vol.getHeader());
  const int nf = vol.getNFrames();
  const size_t bytes =
    sizeof(float) * static_cast<size_t>(vol.getNVoxelsPerFrame()) * nf;
  device::Buffer img(bytes);
  img.upload(vol.getAllFrames(), bytes);
  device::check(
    fir_convolve(device::context(), &v, img.f(), nf, fwhm.data(), cutRadius));
  img.download(vol.getAllFrames(), bytes);
  vol.setActiveFrame(nf - 1);
}

void cutCircle(VolData& vol, float cutRadius)
{
  if (!(cutRadius > 0.0)) return;
  const fir_vol_desc v = device::describe(vol.getHeader());
  const size_t bytes = sizeof(float) * vol.getNVoxelsPerFrame();
  device::Buffer img(bytes);
  img.upload(vol.getDataArray(), bytes);
  device::check(fir_cut_circle(device::context(), &v, img.f(), 1, cutRadius));
  img.download(vol.getDataArray(), bytes);
}

void applyMask(VolData& vol, VolData& mask)
{
  if (vol.getHeader() != mask.getHeader()) error("Mask volume not the same size");
  float* a = vol.getDataArray();
  const float* m = mask.getDataArray();
  for (int i = 0; i < vol.getNVoxelsPerFrame(); ++i)
    if (m[i] <= 0) a[i] = 0.0f;
}

void HounsfieldToMuMap(VolData& vol)
{
  device::huToMu(vol.getDataArray(), vol.getNVoxelsPerFrame());
}
}
