// pm_image.cpp -- image output and comparison for the photon mapper.
//
// The reference writes only a uchar4 frame into a GL buffer
// (photonMappingKernel.cu:1451-1459) and has no way to save or compare
// frames.  This tool renders a preset through the pm_* C ABI into FP32 and
// writes it as PFM (and optionally PPM), and compares two PFM images with the
// metrics used for the precision report.
//
//   pm_image render <preset> <out.pfm> [--ppm out.ppm] [--size WxH] [--spp N]
//                   [--cache FRAMES]   (learned-cache render after FRAMES training frames)
//   pm_image diff <a.pfm> <ref.pfm> [--exposure E]
//     prints: rmse_rel display_rmse median_rel p999_rel max_rel
//     rmse_rel     = sqrt(mean (a - ref)^2) / mean(ref)
//     display_rmse = the same on min(1, max(0, x * E)), relative to the mean display value
//     *_rel        = |a - ref| / ref per value; max over values above 1% of mean(ref)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../../include/photon_mapper.h"

namespace {

bool write_pfm(const char* path, int w, int h, const std::vector<float>& rgb) {
  FILE* fp = fopen(path, "wb");
  if (!fp) return false;
  fprintf(fp, "PF\n%d %d\n-1.0\n", w, h);  // negative scale: little-endian
  for (int y = h - 1; y >= 0; --y) fwrite(&rgb[(size_t)y * w * 3], sizeof(float), (size_t)w * 3, fp);  // bottom row first
  return fclose(fp) == 0;
}

bool read_pfm(const char* path, int* w, int* h, std::vector<float>& rgb) {
  FILE* fp = fopen(path, "rb");
  if (!fp) return false;
  char tag[3] = {0, 0, 0};
  float scale = 0.0f;
  bool ok = fscanf(fp, "%2s %d %d %f", tag, w, h, &scale) == 4 && !strcmp(tag, "PF") && scale < 0.0f && *w > 0 &&
            *h > 0;
  if (ok) ok = fgetc(fp) != EOF;  // the single whitespace byte after the scale
  if (ok) {
    rgb.assign((size_t)*w * *h * 3, 0.0f);
    for (int y = *h - 1; y >= 0 && ok; --y)
      ok = fread(&rgb[(size_t)y * *w * 3], sizeof(float), (size_t)*w * 3, fp) == (size_t)*w * 3;
  }
  fclose(fp);
  return ok;
}

bool write_ppm(const char* path, int w, int h, const std::vector<pm_uchar4>& px) {
  FILE* fp = fopen(path, "wb");
  if (!fp) return false;
  fprintf(fp, "P6\n%d %d\n255\n", w, h);
  for (size_t i = 0; i < px.size(); ++i) {
    unsigned char c[3] = {px[i].x, px[i].y, px[i].z};
    fwrite(c, 1, 3, fp);
  }
  return fclose(fp) == 0;
}

int fail(const char* what, int st) {
  fprintf(stderr, "pm_image: %s failed: %s\n", what, pm_status_string(st));
  return 1;
}

int render(int argc, char** argv) {
  if (argc < 4) return 2;
  const int preset = atoi(argv[2]);
  const char* out = argv[3];
  const char* ppm = nullptr;
  int w = 0, h = 0, spp = 0, cache_frames = -1;
  for (int i = 4; i < argc; ++i) {
    if (!strcmp(argv[i], "--ppm") && i + 1 < argc) ppm = argv[++i];
    else if (!strcmp(argv[i], "--size") && i + 1 < argc) sscanf(argv[++i], "%dx%d", &w, &h);
    else if (!strcmp(argv[i], "--spp") && i + 1 < argc) spp = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--cache") && i + 1 < argc) cache_frames = atoi(argv[++i]);
  }
  pm_config cfg;
  int st = pm_config_preset(preset, &cfg);
  if (st != PM_OK) return fail("pm_config_preset", st);
  if (w > 0 && h > 0) {
    cfg.width = w;
    cfg.height = h;
  }
  if (spp > 0) cfg.spp = spp;
  pm_context* ctx = nullptr;
  if ((st = pm_create(&cfg, &ctx)) != PM_OK) return fail("pm_create", st);
  const size_t npx = (size_t)cfg.width * cfg.height;
  float* rgb_dev = nullptr;
  pm_uchar4* u8_dev = nullptr;
  if (hipMalloc((void**)&rgb_dev, 12 * npx) != hipSuccess || hipMalloc((void**)&u8_dev, 4 * npx) != hipSuccess) {
    fprintf(stderr, "pm_image: hipMalloc failed\n");
    return 1;
  }
  if ((st = pm_emit(ctx)) != PM_OK) return fail("pm_emit", st);
  if ((st = pm_build(ctx)) != PM_OK) return fail("pm_build", st);
  if (cache_frames >= 0) {
    for (int f = 0; f <= cache_frames && st == PM_OK; ++f)
      st = pm_render_cache_crop(ctx, 0, 0, cfg.width, cfg.height, rgb_dev, u8_dev);
    if (st != PM_OK) return fail("pm_render_cache_crop", st);
  } else if ((st = pm_render(ctx, rgb_dev, u8_dev)) != PM_OK) {
    return fail("pm_render", st);
  }
  std::vector<float> rgb(3 * npx);
  std::vector<pm_uchar4> px(npx);
  if (hipMemcpy(rgb.data(), rgb_dev, 12 * npx, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(px.data(), u8_dev, 4 * npx, hipMemcpyDeviceToHost) != hipSuccess)
    return 1;
  (void)hipFree(rgb_dev);
  (void)hipFree(u8_dev);
  pm_destroy(ctx);
  if (!write_pfm(out, cfg.width, cfg.height, rgb)) return 1;
  if (ppm && !write_ppm(ppm, cfg.width, cfg.height, px)) return 1;
  printf("wrote %s (%dx%d, %d spp%s)\n", out, cfg.width, cfg.height, cfg.spp, cache_frames >= 0 ? ", cache" : "");
  return 0;
}

int diff(int argc, char** argv) {
  if (argc < 4) return 2;
  float exposure = 0.1f;
  for (int i = 4; i < argc; ++i)
    if (!strcmp(argv[i], "--exposure") && i + 1 < argc) exposure = (float)atof(argv[++i]);
  int wa, ha, wb, hb;
  std::vector<float> a, b;
  if (!read_pfm(argv[2], &wa, &ha, a) || !read_pfm(argv[3], &wb, &hb, b)) {
    fprintf(stderr, "pm_image: cannot read the PFM inputs\n");
    return 1;
  }
  if (wa != wb || ha != hb) {
    fprintf(stderr, "pm_image: sizes differ (%dx%d vs %dx%d)\n", wa, ha, wb, hb);
    return 1;
  }
  const size_t n = a.size();
  double se = 0.0, sref = 0.0, sed = 0.0, sdref = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double d = (double)a[i] - (double)b[i];
    se += d * d;
    sref += fabs((double)b[i]);
    const double da = std::min(1.0, std::max(0.0, (double)a[i] * exposure));
    const double db = std::min(1.0, std::max(0.0, (double)b[i] * exposure));
    sed += (da - db) * (da - db);
    sdref += db;
  }
  const double mean_ref = sref / (double)n;
  std::vector<double> rel(n);
  double max_rel = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double ref = fabs((double)b[i]);
    rel[i] = fabs((double)a[i] - (double)b[i]) / std::max(ref, 1e-30);
    if (ref > 0.01 * mean_ref) max_rel = std::max(max_rel, rel[i]);
  }
  std::sort(rel.begin(), rel.end());
  const double median = n % 2 ? rel[n / 2] : 0.5 * (rel[n / 2 - 1] + rel[n / 2]);
  // percentile with linear interpolation between closest ranks
  const double pos = 0.999 * (double)(n - 1);
  const size_t lo = (size_t)pos;
  const double p999 = lo + 1 < n ? rel[lo] + (pos - (double)lo) * (rel[lo + 1] - rel[lo]) : rel[lo];
  printf("rmse_rel %.6e display_rmse %.6e median_rel %.6e p999_rel %.6e max_rel %.6e\n",
         sqrt(se / (double)n) / mean_ref, sqrt(sed / (double)n) / (sdref / (double)n), median, p999, max_rel);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 2 && !strcmp(argv[1], "render")) return render(argc, argv);
  if (argc >= 2 && !strcmp(argv[1], "diff")) return diff(argc, argv);
  fprintf(stderr,
          "usage: pm_image render <preset> <out.pfm> [--ppm out.ppm] [--size WxH] [--spp N] [--cache FRAMES]\n"
          "       pm_image diff <a.pfm> <ref.pfm> [--exposure E]\n");
  return 2;
}
