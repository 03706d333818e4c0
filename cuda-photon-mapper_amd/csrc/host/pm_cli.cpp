// pm_cli.cpp -- headless host driver: the reference's GLUT/PBO host
// (simpleGLMain.cpp:64-88, simplePBO.cpp:125-182, callbacksPBO.cpp:47-101)
// reduced to its call order, with the GL pixel buffer replaced by a device
// buffer from hipMalloc.  Per frame: init random numbers once, emit photons
// (when emission is on), render; the frame is written as a binary PPM.
//
//   pm_cli [out.ppm] [frames] [width] [height] [--media] [--interp] [--anim]
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../../include/photon_mapper.h"

int main(int argc, char** argv) {
  const char* out = argc > 1 ? argv[1] : "frame.ppm";
  int frames = argc > 2 ? atoi(argv[2]) : 1;
  unsigned w = argc > 3 ? (unsigned)atoi(argv[3]) : 512u;
  unsigned h = argc > 4 ? (unsigned)atoi(argv[4]) : 512u;
  bool media = false, interp = false, anim = false;
  for (int i = 5; i < argc; ++i) {
    if (!strcmp(argv[i], "--media")) media = true;
    if (!strcmp(argv[i], "--interp")) interp = true;
    if (!strcmp(argv[i], "--anim")) anim = true;
  }
  pm_uchar4* pbo = nullptr;
  if (hipMalloc((void**)&pbo, sizeof(pm_uchar4) * w * h) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  float animTime = 0.0f;
  launch_init_random_numbers_kernel();
  for (int f = 0; f < frames; ++f) {
    launch_emit_photons_kernel(pbo, w, h, animTime, interp, media);
    launch_photon_mapping_kernel(pbo, w, h, animTime, interp, media);
    if (anim) animTime += 0.1f;
  }
  std::vector<pm_uchar4> px((size_t)w * h);
  if (hipMemcpy(px.data(), pbo, sizeof(pm_uchar4) * w * h, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  FILE* fp = fopen(out, "wb");
  if (!fp) return 1;
  fprintf(fp, "P6\n%u %u\n255\n", w, h);
  for (size_t i = 0; i < px.size(); ++i) {
    unsigned char rgb[3] = {px[i].x, px[i].y, px[i].z};
    fwrite(rgb, 1, 3, fp);
  }
  fclose(fp);
  (void)hipFree(pbo);
  printf("wrote %s (%ux%u, %d frame(s))\n", out, w, h, frames);
  return 0;
}
