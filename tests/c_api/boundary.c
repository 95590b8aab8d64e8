/* A C application of the Open VKL API, built against include/ and linked to
 * libopenvkl_hip.so, exactly as an existing Open VKL user switches devices
 * (examples/vklTutorialCPU.c:17-290 of the reference follows the same
 * sequence). It builds a structuredRegular volume and a vdb volume from
 * integer-valued procedural data, samples and differentiates them at points
 * from a 32-bit LCG (host pointers), iterates intervals and hits with the
 * per-ray calls, and writes every result to the binary file argv[1]:
 *
 *   float dense_samples[NPTS]
 *   float dense_gradients[NGRAD][3]
 *   float vdb_samples[NPTS]
 *   per ray (NRAYS): uint32 n, float intervals[n][5]
 *   per ray (NRAYS): uint32 n, float hits[n][3]
 *
 * The expected file content comes from the CPU oracle on the same inputs
 * (tests/golden/make_c_boundary.py). Exit status 0 on success. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <openvkl/device/openvkl.h>
#include <openvkl/openvkl.h>

#define DIM 32
#define NPTS 4096
#define NGRAD 256
#define NRAYS 16

static uint32_t lcg_state = 12345u;
static uint32_t lcg(void)
{
  lcg_state = lcg_state * 1664525u + 1013904223u;
  return lcg_state;
}
/* lo + U[0, 1) * extent with a 24-bit uniform, all in binary32 */
static float uniform(float lo, float extent)
{
  return (float)(lcg() >> 8) * (1.0f / 16777216.0f) * extent + lo;
}

static float dense_value(int x, int y, int z) { return (float)((x * 7 + y * 13 + z * 29) % 97) / 8.0f; }
static float leaf_value(int x, int y, int z) { return (float)((x * 5 + y * 11 + z * 3) % 61) / 16.0f; }

static int failed = 0;
static void on_error(void *user, VKLError code, const char *message)
{
  (void)user;
  fprintf(stderr, "vkl error %d: %s\n", (int)code, message);
  failed = 1;
}

static int write_all(FILE *f, const void *p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes; }

int main(int argc, char **argv)
{
  if (argc != 2) {
    fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 2;
  }
  FILE *out = fopen(argv[1], "wb");
  if (!out) return 2;

  vklInit();
  VKLDevice device = vklNewDevice("hip");
  if (!device) return 3;
  vklDeviceSetErrorCallback(device, on_error, NULL);
  vklCommitDevice(device);

  /* ---- structuredRegular ------------------------------------------------ */
  float *voxels = (float *)malloc(sizeof(float) * DIM * DIM * DIM);
  for (int z = 0; z < DIM; z++)
    for (int y = 0; y < DIM; y++)
      for (int x = 0; x < DIM; x++) voxels[(z * DIM + y) * DIM + x] = dense_value(x, y, z);
  VKLVolume dense = vklNewVolume(device, "structuredRegular");
  vklSetVec3i(dense, "dimensions", DIM, DIM, DIM);
  vklSetVec3f(dense, "gridOrigin", -1.0f, 0.5f, 2.0f);
  vklSetVec3f(dense, "gridSpacing", 0.5f, 1.0f, 1.5f);
  VKLData data = vklNewData(device, (size_t)DIM * DIM * DIM, VKL_FLOAT, voxels, VKL_DATA_DEFAULT, 0);
  vklSetData(dense, "data", data);
  vklRelease(data);
  vklCommit(dense);
  VKLSampler ds = vklNewSampler(dense);
  vklCommit(ds);

  vkl_vec3f *pts = (vkl_vec3f *)malloc(sizeof(vkl_vec3f) * NPTS);
  for (int i = 0; i < NPTS; i++) {
    pts[i].x = uniform(-1.5f, 17.0f);
    pts[i].y = uniform(0.0f, 33.0f);
    pts[i].z = uniform(1.0f, 49.0f);
  }
  float *samples = (float *)malloc(sizeof(float) * NPTS);
  vklComputeSampleN(&ds, NPTS, pts, samples, 0, NULL);
  vkl_vec3f *grads = (vkl_vec3f *)malloc(sizeof(vkl_vec3f) * NGRAD);
  vklComputeGradientN(&ds, NGRAD, pts, grads, 0, NULL);
  if (!write_all(out, samples, sizeof(float) * NPTS) || !write_all(out, grads, sizeof(vkl_vec3f) * NGRAD)) return 4;

  /* ---- vdb: 4 x 4 x 4 leaves of 8^3 voxels, every fifth one missing ---- */
  uint32_t levels[64], formats[64];
  vkl_vec3i origins[64];
  VKLData leafData[64];
  float leaf[512];
  int numLeaves = 0;
  for (int lx = 0; lx < 4; lx++)
    for (int ly = 0; ly < 4; ly++)
      for (int lz = 0; lz < 4; lz++) {
        if ((lx * 16 + ly * 4 + lz) % 5 == 4) continue;
        for (int vx = 0; vx < 8; vx++)
          for (int vy = 0; vy < 8; vy++)
            for (int vz = 0; vz < 8; vz++)
              leaf[vx * 64 + vy * 8 + vz] = leaf_value(lx * 8 + vx, ly * 8 + vy, lz * 8 + vz);
        levels[numLeaves]    = 3;
        formats[numLeaves]   = VKL_FORMAT_DENSE_ZYX;
        origins[numLeaves].x = lx * 8;
        origins[numLeaves].y = ly * 8;
        origins[numLeaves].z = lz * 8;
        leafData[numLeaves]  = vklNewData(device, 512, VKL_FLOAT, leaf, VKL_DATA_DEFAULT, 0);
        numLeaves++;
      }
  VKLVolume vdb = vklNewVolume(device, "vdb");
  VKLData d = vklNewData(device, numLeaves, VKL_UINT, levels, VKL_DATA_DEFAULT, 0);
  vklSetData(vdb, "node.level", d);
  vklRelease(d);
  d = vklNewData(device, numLeaves, VKL_VEC3I, origins, VKL_DATA_DEFAULT, 0);
  vklSetData(vdb, "node.origin", d);
  vklRelease(d);
  d = vklNewData(device, numLeaves, VKL_UINT, formats, VKL_DATA_DEFAULT, 0);
  vklSetData(vdb, "node.format", d);
  vklRelease(d);
  d = vklNewData(device, numLeaves, VKL_DATA, leafData, VKL_DATA_DEFAULT, 0);
  vklSetData(vdb, "node.data", d);
  vklRelease(d);
  for (int i = 0; i < numLeaves; i++) vklRelease(leafData[i]);
  vklCommit(vdb);
  VKLSampler vs = vklNewSampler(vdb);
  vklCommit(vs);
  for (int i = 0; i < NPTS; i++) {
    pts[i].x = uniform(-1.0f, 34.0f);
    pts[i].y = uniform(-1.0f, 34.0f);
    pts[i].z = uniform(-1.0f, 34.0f);
  }
  vklComputeSampleN(&vs, NPTS, pts, samples, 0, NULL);
  if (!write_all(out, samples, sizeof(float) * NPTS)) return 4;

  /* ---- per-ray interval and hit iteration on the vdb volume ------------- */
  VKLIntervalIteratorContext ictx = vklNewIntervalIteratorContext(vs);
  vklCommit(ictx);
  VKLHitIteratorContext hctx = vklNewHitIteratorContext(vs);
  float iso[2] = {1.0f, 2.5f};
  d = vklNewData(device, 2, VKL_FLOAT, iso, VKL_DATA_DEFAULT, 0);
  vklSetData(hctx, "values", d);
  vklRelease(d);
  vklCommit(hctx);
  vkl_vec3f org[NRAYS], dir[NRAYS];
  for (int r = 0; r < NRAYS; r++) {
    org[r].x = uniform(-8.0f, 4.0f);
    org[r].y = uniform(0.0f, 32.0f);
    org[r].z = uniform(0.0f, 32.0f);
    dir[r].x = 1.0f;
    dir[r].y = uniform(-0.5f, 1.0f);
    dir[r].z = uniform(-0.5f, 1.0f);
  }
  const vkl_range1f tRange = {0.0f, INFINITY};
  char *ibuf = (char *)malloc(vklGetIntervalIteratorSize(&ictx));
  for (int r = 0; r < NRAYS; r++) {
    VKLIntervalIterator it = vklInitIntervalIterator(&ictx, &org[r], &dir[r], &tRange, 0.0f, ibuf);
    float iv[256][5];
    uint32_t n = 0;
    VKLInterval interval;
    while (n < 256 && vklIterateInterval(it, &interval)) {
      iv[n][0] = interval.tRange.lower;
      iv[n][1] = interval.tRange.upper;
      iv[n][2] = interval.valueRange.lower;
      iv[n][3] = interval.valueRange.upper;
      iv[n][4] = interval.nominalDeltaT;
      n++;
    }
    if (!write_all(out, &n, 4) || !write_all(out, iv, sizeof(float) * 5 * n)) return 4;
  }
  char *hbuf = (char *)malloc(vklGetHitIteratorSize(&hctx));
  for (int r = 0; r < NRAYS; r++) {
    VKLHitIterator it = vklInitHitIterator(&hctx, &org[r], &dir[r], &tRange, 0.0f, hbuf);
    float hv[64][3];
    uint32_t n = 0;
    VKLHit hit;
    while (n < 64 && vklIterateHit(it, &hit)) {
      hv[n][0] = hit.t;
      hv[n][1] = hit.sample;
      hv[n][2] = hit.epsilon;
      n++;
    }
    if (!write_all(out, &n, 4) || !write_all(out, hv, sizeof(float) * 3 * n)) return 4;
  }

  free(ibuf);
  free(hbuf);
  vklRelease(hctx);
  vklRelease(ictx);
  vklRelease(vs);
  vklRelease(vdb);
  vklRelease(ds);
  vklRelease(dense);
  vklReleaseDevice(device);
  free(pts);
  free(samples);
  free(grads);
  free(voxels);
  fclose(out);
  return failed ? 5 : 0;
}
