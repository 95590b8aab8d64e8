// Observers: "InnerNode" on vdb volumes and "LeafNodeAccess" on their
// samplers (openvkl/include/openvkl/observer.h:21-75).
//
// InnerNode walks the host mirror of the node table; LeafNodeAccess reads the
// sampler's device access-stamp buffer, which the sparse traversal writes
// (SparseSource::traverse).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "objects.h"

namespace vklhip {

// ---- InnerNode ---------------------------------------------------------------------

InnerNodeObserver::InnerNodeObserver(Device *d, Volume *v) : Observer(d), volume(v) { volume->retain(); }

InnerNodeObserver::~InnerNodeObserver() { volume->release(); }

// VdbInnerNodeObserver.cpp:96-186. Voxels of levels 0..maxDepth in the order
// level, node, x, y, z (the reference fills its buffer in parallel, so its
// order is unspecified); a voxel is reported when it is a leaf pointer, or on
// level maxDepth when it is not empty. maxDepth is clamped to the deepest
// level that holds voxels (2): the reference's bound of 3 names a level
// without a voxel array, and every leaf pointer is reported by depth 2.
void InnerNodeObserver::commit()
{
  const int requested = getInt("maxDepth", 1);
  const uint32_t maxDepth = std::min<uint32_t>((uint32_t)requested, 2u);
  const uint32_t nattr    = volume->numAttributes();
  numFloats               = 6 + 2 * nattr;
  nodes.clear();

  // Node origins (root-relative) per level, from the child pointers.
  std::vector<uint32_t> origins[3];
  origins[0].assign(3, 0u);
  for (uint32_t l = 0; l < 2; l++) {
    const uint64_t numNodes = volume->grid.levels[l + 1].numNodes;
    origins[l + 1].assign(3 * numNodes, 0u);
    const std::vector<uint64_t> &vox = volume->hostVoxels[l];
    const uint32_t logRes = kLogRes[l], ext = kRes[l + 1];
    for (uint64_t i = 0; i < vox.size(); i++) {
      if ((vox[i] & 3u) != 2u) continue;
      const uint64_t node = i >> (3 * logRes), local = i & ((1ull << (3 * logRes)) - 1);
      const uint32_t vx = (uint32_t)(local >> (2 * logRes)), vy = (uint32_t)(local >> logRes) & ((1u << logRes) - 1),
                     vz = (uint32_t)local & ((1u << logRes) - 1);
      const uint64_t child = vox[i] >> 2;
      origins[l + 1][3 * child]     = origins[l][3 * node] + vx * ext;
      origins[l + 1][3 * child + 1] = origins[l][3 * node + 1] + vy * ext;
      origins[l + 1][3 * child + 2] = origins[l][3 * node + 2] + vz * ext;
    }
  }

  for (uint32_t l = 0; l <= maxDepth; l++) {
    const std::vector<uint64_t> &vox = volume->hostVoxels[l];
    const std::vector<Range1f> &rng  = volume->hostRanges[l];
    const uint32_t logRes = kLogRes[l], ext = kRes[l + 1];
    for (uint64_t i = 0; i < vox.size(); i++) {
      const uint64_t v = vox[i];
      const uint32_t tag = (uint32_t)v & 3u;
      if (!(tag == 3u || (l == maxDepth && tag != 0u))) continue;
      const uint64_t node = i >> (3 * logRes), local = i & ((1ull << (3 * logRes)) - 1);
      const uint32_t vx = (uint32_t)(local >> (2 * logRes)), vy = (uint32_t)(local >> logRes) & ((1u << logRes) - 1),
                     vz = (uint32_t)local & ((1u << logRes) - 1);
      float box[6];
      const uint32_t off[3] = {vx * ext, vy * ext, vz * ext};
      for (int c = 0; c < 3; c++) {
        const int32_t lo = volume->grid.rootOrigin[c] + (int32_t)(origins[l][3 * node + c] + off[c]);
        box[c]     = (float)lo;
        box[3 + c] = (float)lo + (float)ext;
      }
      float obj[6];
      transformBounds(volume->i2o, box, obj);
      const size_t at = nodes.size();
      nodes.resize(at + numFloats);
      std::memcpy(&nodes[at], obj, sizeof(obj));
      for (uint32_t a = 0; a < nattr; a++) {
        nodes[at + 6 + 2 * a] = rng[i * nattr + a].lower;
        nodes[at + 7 + 2 * a] = rng[i * nattr + a].upper;
      }
    }
  }
  built = true;
}

const void *InnerNodeObserver::map()
{
  if (!built) commit();
  return nodes.data();
}

// ---- LeafNodeAccess -------------------------------------------------------------

LeafAccessObserver::LeafAccessObserver(Device *d, Sampler *s) : Observer(d), sampler(s)
{
  sampler->retain();
  const uint64_t n = sampler->volume->numLeaves;
  device->makeCurrent();
  if (!sampler->accessBuf.ptr && n) {
    sampler->accessBuf.allocate(n * sizeof(uint32_t));
    checkHip(hipMemsetAsync(sampler->accessBuf.ptr, 0, n * sizeof(uint32_t), device->stream), "hipMemsetAsync");
  }
  epoch = ++sampler->accessEpoch;
  sampler->liveObservers++;
  sampler->applyObservation();
  snapshot.assign(n, 0u);
}

LeafAccessObserver::~LeafAccessObserver()
{
  sampler->liveObservers--;
  try {
    sampler->applyObservation();
  } catch (...) {
  }
  sampler->release();
}

// The reference hands out its live buffer; here the device stamps are read
// once per map (after the device's stream drains), entry i = 1 when leaf i
// was reached at or after this observer's epoch.
const void *LeafAccessObserver::map()
{
  const uint64_t n = snapshot.size();
  if (n) {
    device->makeCurrent();
    checkHip(hipStreamSynchronize(device->stream), "hipStreamSynchronize");
    checkHip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    checkHip(hipMemcpy(snapshot.data(), sampler->accessBuf.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost),
             "hipMemcpy");
    for (uint64_t i = 0; i < n; i++) snapshot[i] = snapshot[i] >= epoch ? 1u : 0u;
  }
  return snapshot.data();
}

}  // namespace vklhip
