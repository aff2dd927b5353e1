// PLY point-cloud export of raycast maps: the step after the raycast that the
// reference does with PCL (ThreeDViewer::SetUpPointClouds,
// threeD_viewer.cpp:20-55): every pixel whose vertex z is non-zero becomes a
// point; colour, when given, is read as the reference reads its colour image
// (3 bytes per pixel in B, G, R order).
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/kinfu_amd.h"

extern "C" int kf_write_ply(const char* path, int64_t n, const float* x, const float* y,
                            const float* z, const float* nx, const float* ny, const float* nz,
                            const uint8_t* bgr, int32_t binary) {
  if (!path || n < 0 || !x || !y || !z) return KF_ERR_INVALID;
  const bool normals = nx && ny && nz;
  int64_t count = 0;
  for (int64_t i = 0; i < n; ++i) count += z[i] != 0.0f;
  FILE* f = fopen(path, binary ? "wb" : "w");
  if (!f) return KF_ERR_IO;
  fprintf(f, "ply\nformat %s 1.0\n", binary ? "binary_little_endian" : "ascii");
  fprintf(f, "element vertex %lld\n", (long long)count);
  fprintf(f, "property float x\nproperty float y\nproperty float z\n");
  if (normals) fprintf(f, "property float nx\nproperty float ny\nproperty float nz\n");
  if (bgr) fprintf(f, "property uchar red\nproperty uchar green\nproperty uchar blue\n");
  fprintf(f, "end_header\n");
  std::vector<uint8_t> rec;
  for (int64_t i = 0; i < n; ++i) {
    if (z[i] == 0.0f) continue;
    const float v[6] = {x[i], y[i], z[i], normals ? nx[i] : 0.f, normals ? ny[i] : 0.f,
                        normals ? nz[i] : 0.f};
    const uint8_t c[3] = {bgr ? bgr[3 * i + 2] : (uint8_t)0, bgr ? bgr[3 * i + 1] : (uint8_t)0,
                          bgr ? bgr[3 * i] : (uint8_t)0};
    if (binary) {
      rec.clear();
      const int nf = normals ? 6 : 3;
      rec.resize(nf * 4 + (bgr ? 3 : 0));
      memcpy(rec.data(), v, nf * 4);  // little-endian host
      if (bgr) memcpy(rec.data() + nf * 4, c, 3);
      if (fwrite(rec.data(), 1, rec.size(), f) != rec.size()) {
        fclose(f);
        return KF_ERR_IO;
      }
    } else {
      fprintf(f, "%.9g %.9g %.9g", v[0], v[1], v[2]);
      if (normals) fprintf(f, " %.9g %.9g %.9g", v[3], v[4], v[5]);
      if (bgr) fprintf(f, " %u %u %u", c[0], c[1], c[2]);
      fprintf(f, "\n");
    }
  }
  return fclose(f) == 0 ? KF_OK : KF_ERR_IO;
}
