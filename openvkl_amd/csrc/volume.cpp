// Volume commit: parameter parsing, VDB node-table build, device upload and
// the majorant (value range) build for "structuredRegular" and "vdb".
//
// The node table follows openvkl/devices/cpu/volume/vdb/VdbVolume.cpp:71-392,
// 583-1062: leaves binned by level, inner nodes numbered in first-encounter
// order, voxel codes empty 00 / child idx<<2|10 / leaf idx<<6|TF<<4|DF<<2|11
// (VdbGrid.h:91-203), value ranges of each leaf's filter-support window
// extended up the parent chain. Per-leaf ranges are computed on the GPU.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <sstream>

#include "kernels.h"
#include "objects.h"

namespace vklhip {

namespace {

const float kInf = std::numeric_limits<float>::infinity();

uint64_t linearIndex(uint32_t l, uint32_t x, uint32_t y, uint32_t z)
{
  const uint32_t m = kRes[l] - 1, s = kTotalLogRes[l + 1], L = kLogRes[l];
  return ((uint64_t)((x & m) >> s) << (2 * L)) + ((uint64_t)((y & m) >> s) << L) + ((z & m) >> s);
}

inline uint64_t childCode(uint64_t idx) { return (idx << 2) | 2u; }
// fmt: data format in bits 0-1, temporal format in bits 2-3.
inline uint64_t leafCode(uint64_t idx, uint32_t fmt) { return (idx << 6) + (((uint64_t)fmt & 15u) << 2) + 3u; }

struct NodeTable
{
  uint64_t numNodes[3] = {0, 0, 0};
  std::vector<uint64_t> voxels[3];
};

// Two passes so the (large) level-2 array is allocated once: pass 1 builds
// levels 0 and 1 and numbers level-2 nodes; pass 2 fills level 2. The node
// numbering equals the reference's single pass (same leaf order).
void buildNodeTable(const std::vector<uint32_t> &level, const std::vector<uint32_t> &format,
                    const std::vector<uint32_t> &offs, const std::vector<uint64_t> &leafId, NodeTable &t)
{
  const uint64_t n = level.size();
  std::vector<std::vector<uint64_t>> binned(kNumLevels);
  for (uint64_t i = 0; i < n; i++) binned[level[i]].push_back(i);

  t.numNodes[0] = 1;
  t.voxels[0].assign(kNumVoxels[0], 0);
  auto insertError = [](uint32_t l) {
    std::ostringstream ss;
    ss << "Attempted to insert a leaf node into a leaf node (level " << l + 1 << ")";
    throw VklError(ss.str());
  };

  // Pass 1: levels 1..3 leaves, descending through levels 0 and 1 only.
  for (uint32_t ll = 1; ll < (uint32_t)kNumLevels; ll++) {
    for (uint64_t i : binned[ll]) {
      const uint32_t x = offs[3 * i], y = offs[3 * i + 1], z = offs[3 * i + 2];
      uint64_t node = 0;
      for (uint32_t l = 0; l < std::min<uint32_t>(ll, 2); l++) {
        uint64_t &v = t.voxels[l][node * kNumVoxels[l] + linearIndex(l, x, y, z)];
        if ((v & 3u) == 3u) insertError(l);
        if ((v & 3u) == 0u) {
          if (l + 1 < ll) {
            const uint64_t child = t.numNodes[l + 1]++;
            v                    = childCode(child);
            if (l + 1 == 1) t.voxels[1].resize(t.numNodes[1] * kNumVoxels[1], 0);
            node = child;
          } else {
            v = leafCode(leafId[i], format[i]);
          }
        } else {
          node = v >> 2;
        }
      }
    }
  }
  // Pass 2: level-3 leaves into level 2.
  t.voxels[2].assign(t.numNodes[2] * kNumVoxels[2], 0);
  for (uint64_t i : binned[3]) {
    const uint32_t x = offs[3 * i], y = offs[3 * i + 1], z = offs[3 * i + 2];
    const uint64_t c0 = t.voxels[0][linearIndex(0, x, y, z)];
    const uint64_t c1 = t.voxels[1][(c0 >> 2) * kNumVoxels[1] + linearIndex(1, x, y, z)];
    uint64_t &v       = t.voxels[2][(c1 >> 2) * kNumVoxels[2] + linearIndex(2, x, y, z)];
    if ((v & 3u) != 0u) insertError(2);
    v = leafCode(leafId[i], format[i]);
  }
}

// rkcommon AffineSpace3f inverse used by DenseVdbVolume.cpp:41-64 and
// VdbVolume.cpp:396-405: l^-1 = adjoint * (1 / det), p' = -(l^-1 * p).
// Input column-major (vx, vy, vz, p); outputs row-major 3x4 buffers as
// writeTransform (VdbVolume.h:63-78).
void buildTransforms(const float c[12], float i2o[12], float o2i[12])
{
  const float vx[3] = {c[0], c[1], c[2]}, vy[3] = {c[3], c[4], c[5]}, vz[3] = {c[6], c[7], c[8]};
  i2o[0] = vx[0]; i2o[1] = vy[0]; i2o[2] = vz[0];
  i2o[3] = vx[1]; i2o[4] = vy[1]; i2o[5] = vz[1];
  i2o[6] = vx[2]; i2o[7] = vy[2]; i2o[8] = vz[2];
  i2o[9] = c[9]; i2o[10] = c[10]; i2o[11] = c[11];
  auto cross = [](const float *a, const float *b, float *r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
  };
  float r0[3], r1[3], r2[3];
  cross(vy, vz, r0);
  cross(vz, vx, r1);
  cross(vx, vy, r2);
  const float det  = vx[0] * r0[0] + vx[1] * r0[1] + vx[2] * r0[2];
  const float rdet = 1.f / det;
  for (int k = 0; k < 3; k++) {
    o2i[k]     = r0[k] * rdet;
    o2i[3 + k] = r1[k] * rdet;
    o2i[6 + k] = r2[k] * rdet;
  }
  const float px = c[9], py = c[10], pz = c[11];
  const float lx = o2i[0] * px + o2i[1] * py + o2i[2] * pz;
  const float ly = o2i[3] * px + o2i[4] * py + o2i[5] * pz;
  const float lz = o2i[6] * px + o2i[7] * py + o2i[8] * pz;
  o2i[9]  = -lx;
  o2i[10] = -ly;
  o2i[11] = -lz;
}

float undefinedBackground()
{
  const uint32_t bits = 0xFFC068B5u;
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

std::vector<float> parseBackground(const ManagedObject &o, uint32_t numAttributes)
{
  std::vector<float> bg(numAttributes, undefinedBackground());
  const ManagedObject::Param *p = o.findParam("background");
  if (!p) return bg;
  if (p->type == VKL_FLOAT) {
    float v;
    std::memcpy(&v, p->bytes.data(), 4);
    std::fill(bg.begin(), bg.end(), v);
  } else if (p->type == VKL_DATA && p->object) {
    const Data *d = static_cast<const Data *>(p->object);
    std::vector<float> v = d->hostArray<float>();
    if (v.size() == 1)
      std::fill(bg.begin(), bg.end(), v[0]);
    else if (v.size() == numAttributes)
      bg = v;
    else
      throw VklError("background must have one value or one per attribute");
  }
  return bg;
}

// verifyTemporalData, temporal_data_verification.h:17-181: the number of
// elements each attribute must hold for this temporal configuration.
uint64_t verifyTemporalData(Device &device, uint64_t numVoxels, uint32_t format, int numTimesteps, Data *indices,
                            Data *times)
{
  if (format == VKL_TEMPORAL_FORMAT_CONSTANT) {
    if (numTimesteps != 0)
      throw VklError("Temporally constant volumes should not have temporallyStructuredNumTimesteps provided");
    if (indices || times)
      throw VklError("Temporally constant volumes should not have temporallyUnstructuredIndices or "
                     "temporallyUnstructuredTimes provided");
    return numVoxels;
  }
  if (format == VKL_TEMPORAL_FORMAT_STRUCTURED) {
    if (indices || times)
      throw VklError("Temporally structured volumes should not have temporallyUnstructuredIndices or "
                     "temporallyUnstructuredTimes provided");
    if (numTimesteps <= 0)
      throw VklError("Temporal format is VKL_TEMPORAL_FORMAT_STRUCTURED but temporallyStructuredNumTimesteps is not "
                     "greater than 0.");
    return numVoxels * (uint64_t)numTimesteps;
  }
  if (numTimesteps != 0)
    throw VklError("Temporally unstructured volumes should not have temporallyStructuredNumTimesteps provided");
  if (!indices)
    throw VklError("temporal format is VKL_TEMPORAL_FORMAT_UNSTRUCTURED but temporallyUnstructuredIndices is not "
                   "specified.");
  if (!times)
    throw VklError("temporal format is VKL_TEMPORAL_FORMAT_UNSTRUCTURED but temporallyUnstructuredTimes is not "
                   "specified.");
  if (indices->numItems < numVoxels + 1) throw VklError("temporallyUnstructuredIndices is too small.");
  const bool need64 = numVoxels >= (uint64_t)std::numeric_limits<uint32_t>::max();
  if (need64 && indices->dataType != VKL_ULONG)
    throw VklError("temporallyUnstructuredIndices must be VKL_ULONG due to attribute data size");
  if (!need64 && indices->dataType == VKL_ULONG)
    device.log(VKL_LOG_WARNING, "WARNING: temporallyUnstructuredIndices is VKL_ULONG when VKL_UINT is sufficient "
                                "and may be more performant");
  if (indices->dataType != VKL_UINT && indices->dataType != VKL_ULONG)
    throw VklError("temporallyUnstructuredIndices must have type VKL_UINT or VKL_ULONG");
  if (times->dataType != VKL_FLOAT) throw VklError("temporallyUnstructuredIndices must have type VKL_FLOAT");
  std::vector<uint64_t> idx(indices->numItems);
  if (indices->dataType == VKL_UINT) {
    const std::vector<uint32_t> i32 = indices->hostArray<uint32_t>();
    std::copy(i32.begin(), i32.end(), idx.begin());
  } else {
    idx = indices->hostArray<uint64_t>();
  }
  const std::vector<float> t = times->hostArray<float>();
  uint64_t end = 0;
  for (uint64_t i = 0; i < numVoxels; ++i) {
    const uint64_t b = idx[i];
    end              = idx[i + 1];
    if (!(b < end)) throw VklError("temporallyUnstructuredIndices must increase strictly monotonically.");
    if (!(end <= t.size()))
      throw VklError("Values in temporallyUnstructuredIndices are out of bounds in temporallyUnstructredTimes.");
    for (uint64_t k = b + 1; k < end; ++k)
      if (!(t[k - 1] < t[k]))
        throw VklError("temporallyUnstructuredTimes must increase strictly monotonically for every voxel.");
    if (t[b] < 0.f || t[b] > 1.f || t[end - 1] < 0.f || t[end - 1] > 1.f)
      throw VklError("Values in temporallyUnstructuredTimes must be bounded by 0.0 and 1.0 for every voxel.");
  }
  return end;
}

}  // namespace

void transformBounds(const float i2o[12], const float box[6], float out[6])
{
  out[0] = out[1] = out[2] = kInf;
  out[3] = out[4] = out[5] = -kInf;
  for (int i = 0; i < 8; i++) {
    const float v[3] = {(i & 1) ? box[3] : box[0], (i & 2) ? box[4] : box[1], (i & 4) ? box[5] : box[2]};
    for (int r = 0; r < 3; r++) {
      const float w = i2o[3 * r] * v[0] + i2o[3 * r + 1] * v[1] + i2o[3 * r + 2] * v[2] + i2o[9 + r];
      out[r]        = std::min(out[r], w);
      out[3 + r]    = std::max(out[3 + r], w);
    }
  }
}

DeviceBuffer *Volume::newBuffer(size_t bytes)
{
  buffers.emplace_back(new DeviceBuffer(bytes));
  return buffers.back().get();
}

// VdbVolume.cpp:586-593: filter / gradientFilter / maxSamplingDepth defaults
// that samplers inherit.
void Volume::commitCommonParams()
{
  filter           = getInt("filter", kFilterLinear);
  gradientFilter   = getInt("gradientFilter", filter);
  maxSamplingDepth = std::min(getInt("maxSamplingDepth", 3), 3);
}

namespace {

// Shared tail of both commits: node table upload, per-leaf ranges on the GPU,
// parent-chain extension on the host, level upload, volume value range.
void finishGrid(Volume &v, const std::vector<uint32_t> &level, const std::vector<uint32_t> &format,
                const std::vector<uint32_t> &offs, const std::vector<uint64_t> &leafId, hipStream_t st)
{
  const uint32_t nattr = v.grid.numAttributes;
  NodeTable t;
  buildNodeTable(level, format, offs, leafId, t);

  // Upload voxel codes (needed by the sparse range kernel).
  DeviceBuffer *vox[3], *rng[3];
  for (int l = 0; l < 3; l++) {
    v.grid.levels[l].numNodes = t.numNodes[l];
    vox[l] = v.newBuffer(t.voxels[l].size() * sizeof(uint64_t));
    vox[l]->upload(t.voxels[l].data(), t.voxels[l].size() * sizeof(uint64_t), st);
    v.grid.levels[l].voxels = vox[l]->as<uint64_t>();
    rng[l] = v.newBuffer(t.voxels[l].size() * nattr * sizeof(Range1f));
    v.grid.levels[l].valueRange = rng[l]->as<Range1f>();
  }

  // Per-leaf ranges (VdbSampler.ispc:460-557), one kernel per attribute.
  const uint64_t n = level.size();
  DeviceBuffer dOffs(n * 3 * sizeof(uint32_t));
  dOffs.upload(offs.data(), n * 3 * sizeof(uint32_t), st);
  DeviceBuffer dRanges(n * sizeof(Range1f));
  std::vector<std::vector<Range1f>> leafRanges(nattr, std::vector<Range1f>(n));
  for (uint32_t a = 0; a < nattr; a++) {
    checkHip(launchLeafRanges(v.grid, a, n, dOffs.as<uint32_t>(), dRanges.as<Range1f>(), st), "leaf ranges");
    checkHip(hipMemcpyAsync(leafRanges[a].data(), dRanges.ptr, n * sizeof(Range1f), hipMemcpyDeviceToHost, st),
             "hipMemcpyAsync");
    checkHip(hipStreamSynchronize(st), "leaf ranges");
  }

  // computeValueRanges, VdbVolume.cpp:361-391.
  for (int l = 0; l < 3; l++) v.hostRanges[l].assign(t.voxels[l].size() * nattr, Range1f{kInf, -kInf});
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t x = offs[3 * i], y = offs[3 * i + 1], z = offs[3 * i + 2];
    uint64_t node = 0;
    for (uint32_t l = 0; l < level[i]; l++) {
      const uint64_t vi = node * kNumVoxels[l] + linearIndex(l, x, y, z);
      for (uint32_t a = 0; a < nattr; a++) {
        Range1f &r = v.hostRanges[l][vi * nattr + a];
        const Range1f &s = leafRanges[a][i];
        r.lower = s.lower < r.lower ? s.lower : r.lower;
        r.upper = s.upper > r.upper ? s.upper : r.upper;
      }
      const uint64_t code = t.voxels[l][vi];
      if ((code & 3u) == 3u) break;
      node = code >> 2;
    }
  }
  for (int l = 0; l < 3; l++) {
    rng[l]->upload(v.hostRanges[l].data(), v.hostRanges[l].size() * sizeof(Range1f), st);
    v.hostVoxels[l] = std::move(t.voxels[l]);
  }
  // Volume value range from level 0 (VdbVolume.cpp:1048-1057).
  v.valueRanges.assign(nattr, Range1f{kInf, -kInf});
  for (uint32_t a = 0; a < nattr; a++)
    for (uint64_t i = 0; i < kNumVoxels[0]; i++) {
      const Range1f &r = v.hostRanges[0][i * nattr + a];
      v.valueRanges[a].lower = r.lower < v.valueRanges[a].lower ? r.lower : v.valueRanges[a].lower;
      v.valueRanges[a].upper = r.upper > v.valueRanges[a].upper ? r.upper : v.valueRanges[a].upper;
    }
  v.numLeaves = n;
}

void uploadAttrViews(Volume &v, hipStream_t st)
{
  DeviceBuffer *b = v.newBuffer(v.attrViews.size() * sizeof(AttrView));
  b->upload(v.attrViews.data(), v.attrViews.size() * sizeof(AttrView), st);
  v.grid.attrs = b->as<AttrView>();
}

// Device copy of a Data array (aliases compact device-resident shared data).
const void *deviceData(Volume &v, const Data &d, hipStream_t st)
{
  if (d.onDevice && d.compact()) return d.base;
  std::vector<uint8_t> h = d.hostCopy();
  DeviceBuffer *b        = v.newBuffer(h.size());
  b->upload(h.data(), h.size(), st);
  return b->ptr;
}

}  // namespace

// DenseVdbVolume::commit / parseStructuredVolumeParameters,
// DenseVdbVolume.cpp:20-190.
void DenseVolume::commit()
{
  device->makeCurrent();
  hipStream_t st = device->stream;
  buffers.clear();
  committed = false;
  commitCommonParams();

  int32_t dims[3];
  if (!getVec3i("dimensions", dims)) throw VklError(toString() + ": missing required 'dimensions' parameter");
  if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) throw VklError("Vdb volumes must have at least one leaf node.");
  float origin[3] = {0.f, 0.f, 0.f}, spacing[3] = {1.f, 1.f, 1.f};
  getVec3f("gridOrigin", origin);
  getVec3f("gridSpacing", spacing);
  int32_t indexOrigin[3] = {0, 0, 0};
  getVec3i("indexOrigin", indexOrigin);

  std::vector<Data *> attrs;
  Data *d = getData("data");
  if (d && d->dataType == VKL_DATA) {
    for (size_t i = 0; i < d->numItems; i++) attrs.push_back(d->item(i));
  } else if (d) {
    attrs.push_back(d);
  } else {
    throw VklError(toString() + ": missing required 'data' parameter");
  }
  const uint64_t nvox = (uint64_t)dims[0] * dims[1] * dims[2];
  for (size_t i = 0; i < attrs.size(); i++) {
    const VKLDataType t = attrs[i]->dataType;
    if (t != VKL_UCHAR && t != VKL_SHORT && t != VKL_USHORT && t != VKL_HALF && t != VKL_FLOAT && t != VKL_DOUBLE) {
      std::ostringstream ss;
      ss << toString() << ": unsupported data element type (attribute " << i << ") for 'data' parameter";
      throw VklError(ss.str());
    }
  }
  // Motion blur data (DenseVdbVolume.cpp:148-189): the format follows from
  // which parameters are present; the attribute size from the format.
  const int numTimesteps = getInt("temporallyStructuredNumTimesteps", 0);
  Data *tuIdx            = getData("temporallyUnstructuredIndices");
  Data *tuTimes          = getData("temporallyUnstructuredTimes");
  const uint32_t tformat = numTimesteps > 0 ? VKL_TEMPORAL_FORMAT_STRUCTURED
                           : tuIdx          ? VKL_TEMPORAL_FORMAT_UNSTRUCTURED
                                            : VKL_TEMPORAL_FORMAT_CONSTANT;
  const uint64_t expectedItems = verifyTemporalData(*device, nvox, tformat, numTimesteps, tuIdx, tuTimes);
  for (size_t i = 0; i < attrs.size(); i++)
    if (attrs[i]->numItems != expectedItems) {
      std::ostringstream ss;
      ss << "attribute " << i << " has " << attrs[i]->numItems << " elements, but expected " << expectedItems;
      throw VklError(ss.str());
    }

  float col[12] = {spacing[0], 0.f, 0.f, 0.f, spacing[1], 0.f, 0.f, 0.f, spacing[2], origin[0], origin[1], origin[2]};
  getAffine3f("indexToObject", col);
  buildTransforms(col, i2o, grid.o2i);

  const bool cellCentered = getBool("cellCentered", false);
  grid.dense            = 1;
  grid.constantCellData = cellCentered ? 1 : 0;
  grid.numAttributes    = (uint32_t)attrs.size();
  grid.maxSamplingDepth = (uint32_t)maxSamplingDepth;
  grid.leafSlot         = nullptr;
  for (int i = 0; i < 3; i++) {
    grid.rootOrigin[i] = indexOrigin[i];
    grid.activeSize[i] = (uint32_t)dims[i];
  }
  // VdbVolume.cpp:752-765, 796-806.
  float ibox[6];
  for (int i = 0; i < 3; i++) {
    ibox[i]     = (float)indexOrigin[i];
    ibox[3 + i] = (float)(indexOrigin[i] + dims[i] - (cellCentered ? 0 : 1));
    grid.domainBox[i]     = ibox[i] - (float)indexOrigin[i];
    grid.domainBox[3 + i] = ibox[3 + i] - (float)indexOrigin[i];
  }
  transformBounds(i2o, ibox, bbox);

  // Voxel layout on the device: x-fastest (the caller's order) or bricks of
  // 2^sx x 2^sy x 2^sz voxels ("brickShift", vec3i, each 0..4).
  int32_t shift[3] = {0, 0, 0};
  getVec3i("brickShift", shift);
  for (int i = 0; i < 3; i++)
    if (shift[i] < 0 || shift[i] > 4) throw VklError("brickShift components must be in [0, 4]");
  // Or "cornerPacked" (bool): each cell stores its 8 trilinear corners
  // (8x the memory; one aligned load per trilinear stencil).
  const bool packed = getBool("cornerPacked", false);
  if (packed && (shift[0] | shift[1] | shift[2]))
    throw VklError("cornerPacked and brickShift cannot be combined", VKL_INVALID_ARGUMENT);
  if (tformat != VKL_TEMPORAL_FORMAT_CONSTANT && (packed || (shift[0] | shift[1] | shift[2])))
    throw VklError("cornerPacked and brickShift are not supported for temporal data", VKL_INVALID_ARGUMENT);
  DenseLayout &L = grid.layout;
  L.sx           = (uint32_t)shift[0];
  L.sy           = (uint32_t)shift[1];
  L.sz           = (uint32_t)shift[2];
  L.cornerPacked = packed ? 1u : 0u;
  const uint64_t nbx = ((uint64_t)dims[0] + (1u << L.sx) - 1) >> L.sx;
  const uint64_t nby = ((uint64_t)dims[1] + (1u << L.sy) - 1) >> L.sy;
  const uint64_t nbz = ((uint64_t)dims[2] + (1u << L.sz) - 1) >> L.sz;
  const uint32_t S   = L.sx + L.sy + L.sz;
  L.strideY          = (nbx << S) << (packed ? 3 : 0);
  L.strideZ          = ((nbx * nby) << S) << (packed ? 3 : 0);
  const bool bricked = S != 0;
  const uint32_t udims[3] = {(uint32_t)dims[0], (uint32_t)dims[1], (uint32_t)dims[2]};

  // "cubicRecords" (bool): also store, per voxel, the 16 elements of its
  // 4x4 (y, z) neighbourhood (AttrView::cubic; 16x the voxel memory), which
  // the tricubic stream kernels load as whole stencils.
  // "cubicStencils" (bool): instead, per cell its whole 4x4x4 stencil (64x
  // the voxel memory; a float32 stencil spans exactly two cache lines, a
  // half one one line).
  const bool cubicRecords  = getBool("cubicRecords", false);
  const bool cubicStencils = getBool("cubicStencils", false);
  if ((cubicRecords || cubicStencils) && tformat != VKL_TEMPORAL_FORMAT_CONSTANT)
    throw VklError("cubicRecords and cubicStencils are not supported for temporal data", VKL_INVALID_ARGUMENT);
  if (cubicRecords && cubicStencils)
    throw VklError("cubicRecords and cubicStencils cannot be combined", VKL_INVALID_ARGUMENT);

  const std::vector<float> bg = parseBackground(*this, grid.numAttributes);
  attrViews.clear();
  for (size_t a = 0; a < attrs.size(); a++) {
    AttrView av{};
    const Data &d   = *attrs[a];
    const size_t es = sizeOfType(d.dataType);
    // The caller's x-fastest elements on the device: its own buffer when
    // that is compact device memory, else a temporary upload.
    std::unique_ptr<DeviceBuffer> tmp;
    auto compactSource = [&]() -> const void * {
      if (d.onDevice && d.compact()) return d.base;
      if (!tmp) {
        std::vector<uint8_t> h = d.hostCopy();
        tmp.reset(new DeviceBuffer(h.size()));
        tmp->upload(h.data(), h.size(), st);
      }
      return tmp->ptr;
    };
    if (packed) {
      DeviceBuffer *dst = newBuffer(nvox * 8 * es);
      checkHip(launchPackCorners(compactSource(), dst->ptr, (uint32_t)es, udims, st), "corner packing");
      checkHip(hipStreamSynchronize(st), "corner packing");
      av.data = dst->ptr;
    } else if (bricked) {
      DeviceBuffer *dst = newBuffer((nbx * nby * nbz << S) * es);
      checkHip(launchSwizzle(compactSource(), dst->ptr, (uint32_t)es, udims, L, (nbx * nby * nbz) << S, st),
               "swizzle");
      checkHip(hipStreamSynchronize(st), "swizzle");
      av.data = dst->ptr;
    } else {
      av.data = deviceData(*this, d, st);
    }
    if (cubicRecords || cubicStencils) {
      DeviceBuffer *dst = newBuffer(nvox * (cubicStencils ? 64 : 16) * es);
      checkHip(launchPackCubicRecords(compactSource(), dst->ptr, (uint32_t)es, udims, cubicStencils, st),
               "cubic records");
      checkHip(hipStreamSynchronize(st), "cubic records");
      av.cubic         = dst->ptr;
      av.cubicStencils = cubicStencils ? 1u : 0u;
    }
    av.tile       = nullptr;
    av.type       = (uint32_t)d.dataType;
    av.background = bg[a];
    attrViews.push_back(av);
  }
  uploadAttrViews(*this, st);

  grid.temporalFormat = tformat;
  grid.numTimesteps   = tformat == VKL_TEMPORAL_FORMAT_STRUCTURED ? (uint32_t)numTimesteps : 0u;
  grid.tuIndices      = nullptr;
  grid.tuIndex64      = 0;
  grid.tuTimes        = nullptr;
  if (tformat == VKL_TEMPORAL_FORMAT_UNSTRUCTURED) {
    grid.tuIndices = deviceData(*this, *tuIdx, st);
    grid.tuIndex64 = tuIdx->dataType == VKL_ULONG ? 1u : 0u;
    grid.tuTimes   = static_cast<const float *>(deviceData(*this, *tuTimes, st));
  }

  // One leaf per 8^3 block, x fastest (DenseVdbVolume.cpp:66-119).
  const uint32_t lx = (dims[0] + 7) / 8, ly = (dims[1] + 7) / 8, lz = (dims[2] + 7) / 8;
  const uint64_t n = (uint64_t)lx * ly * lz;
  std::vector<uint32_t> level(n, 3), format(n, VKL_FORMAT_DENSE_ZYX), offs(3 * n);
  std::vector<uint64_t> leafId(n);
  uint64_t i = 0;
  for (uint32_t k = 0; k < lz; k++)
    for (uint32_t j = 0; j < ly; j++)
      for (uint32_t q = 0; q < lx; q++, i++) {
        offs[3 * i]     = q * 8;
        offs[3 * i + 1] = j * 8;
        offs[3 * i + 2] = k * 8;
        leafId[i]       = i;
      }
  finishGrid(*this, level, format, offs, leafId, st);
  committed = true;
}

// VdbVolume::commit, VdbVolume.cpp:583-1062.
void VdbVolume::commit()
{
  device->makeCurrent();
  hipStream_t st = device->stream;
  buffers.clear();
  committed = false;
  commitCommonParams();

  Data *nodeData = getData("node.data");
  Data *packedD  = getData("nodesPackedDense");
  Data *packedT  = getData("nodesPackedTile");
  if (nodeData && packedD) throw VklError("provide either node.data or nodesPackedDense / nodesPackedTile, not both");
  if (!nodeData && !packedD) throw VklError("node.data or nodesPackedDense / nodesPackedTile must be provided");
  Data *levelD  = getData("node.level");
  Data *originD = getData("node.origin");
  Data *formatD = getData("node.format");
  if (!levelD || !originD || !formatD) throw VklError("node.level, node.origin and node.format must be provided");
  const uint64_t n = levelD->numItems;
  if (n == 0) throw VklError("Vdb volumes must have at least one leaf node.");
  std::vector<uint32_t> level  = levelD->hostArray<uint32_t>();
  std::vector<int32_t> origin  = std::vector<int32_t>(3 * n);
  {
    std::vector<uint8_t> b = originD->hostCopy();
    if (originD->numItems != n) throw VklError("node.level, node.origin, node.format, node.temporalFormat, and node.data must all have the same size");
    std::memcpy(origin.data(), b.data(), 12 * n);
  }
  std::vector<uint32_t> format = formatD->hostArray<uint32_t>();
  // Motion blur per node (VdbVolume.cpp:441-457, 710-747): node.temporalFormat
  // (default constant) with per-node numbers of time steps or unstructured
  // index / time arrays; node.data only.
  Data *tfD    = getData("node.temporalFormat");
  Data *stepsD = getData("node.temporallyStructuredNumTimesteps");
  Data *tuIdxD = getData("node.temporallyUnstructuredIndices");
  Data *tuTmD  = getData("node.temporallyUnstructuredTimes");
  std::vector<uint32_t> tformat = tfD ? tfD->hostArray<uint32_t>() : std::vector<uint32_t>(n, 0u);
  if (formatD->numItems != n || (nodeData && nodeData->numItems != n) || tformat.size() != n)
    throw VklError("node.level, node.origin, node.format, node.temporalFormat, and node.data must all have the same size");
  if (stepsD && stepsD->numItems != n)
    throw VklError("If node.temporallyStructuredNumTimesteps is set, it must have the same size as node.data");
  if (tuIdxD && tuIdxD->numItems != n)
    throw VklError("If node.temporallyUnstructuredIndices is set, it must have the same size as node.data");
  if (tuTmD && tuTmD->numItems != n)
    throw VklError("If node.temporallyUnstructuredTimes is set, it must have the same size as node.data");
  const std::vector<int32_t> steps = stepsD ? stepsD->hostArray<int32_t>() : std::vector<int32_t>(n, 0);
  bool anyTemporal = false;
  for (uint32_t f : tformat) anyTemporal |= f != VKL_TEMPORAL_FORMAT_CONSTANT;
  if (anyTemporal && packedD) throw VklError("only temporally constant data is supported for packed dense / tile data");

  // verifyLevel / verifyNodeDataFormat (VdbVolume.cpp:536-563).
  uint64_t numDense = 0, numTile = 0;
  std::vector<uint64_t> slot(n);
  for (uint64_t i = 0; i < n; i++) {
    if (level[i] >= (uint32_t)kNumLevels) {
      std::ostringstream ss;
      ss << "invalid node level " << level[i] << " for this vdb configuration";
      throw VklError(ss.str());
    }
    if (level[i] == 0) throw VklError("there must not be any leaf nodes on level 0");
    if (format[i] == VKL_FORMAT_DENSE_ZYX) {
      if (level[i] + 1 < (uint32_t)kNumLevels) throw VklError("leaf nodes are only supported on the lowest level.");
      slot[i] = numDense++;
    } else if (format[i] == VKL_FORMAT_TILE) {
      slot[i] = numTile++;
    } else {
      throw VklError("invalid format specified");
    }
  }

  // Attribute data.
  std::vector<VKLDataType> types;
  attrViews.clear();
  if (packedD) {
    const uint32_t na = (uint32_t)packedD->numItems;
    for (uint32_t a = 0; a < na; a++) {
      Data *pd = packedD->item(a);
      types.push_back(pd->dataType);
      if (pd->numItems != numDense * 512) throw VklError("nodesPackedDense has incorrect number of elements");
      AttrView av{};
      av.type = (uint32_t)pd->dataType;
      av.data = deviceData(*this, *pd, st);
      if (packedT) {
        Data *pt = packedT->item(a);
        if (pt->dataType != pd->dataType) throw VklError("nodesPackedTile has inconsistent attribute type");
        if (pt->numItems != numTile) throw VklError("nodesPackedTile has incorrect number of elements");
        av.tile = deviceData(*this, *pt, st);
      } else if (numTile) {
        throw VklError("nodesPackedTile has incorrect number of elements");
      }
      attrViews.push_back(av);
    }
    grid.leafSlot = nullptr;
  } else {
    // node.data: per-leaf arrays (or per-leaf VKL_DATA of per-attribute
    // arrays), gathered into packed device arrays plus a leaf -> slot table.
    // With temporal leaves every leaf's elements go, in input order, into one
    // array per attribute (leaf i at leafDataOffset[i]), and the unstructured
    // index / time arrays into two flat arrays.
    Data *first          = nodeData->item(0);
    const bool multi     = first->dataType == VKL_DATA;
    const uint32_t na    = multi ? (uint32_t)first->numItems : 1;
    for (uint32_t a = 0; a < na; a++) types.push_back(multi ? first->item(a)->dataType : first->dataType);
    std::vector<std::vector<uint8_t>> dense(na), tile(na);
    std::vector<uint64_t> dataOff(n), idxOff(n, 0), timesOff(n, 0), flatIdx;
    std::vector<uint32_t> leafSteps(n, 0);
    std::vector<float> flatTimes;
    uint64_t dataTotal = 0;
    for (uint64_t i = 0; i < n; i++) {
      Data *ld = nodeData->item(i);
      if ((ld->dataType == VKL_DATA) != multi) throw VklError("All nodes must have the same VKLDataType in vdb volumes.");
      const uint64_t voxels = format[i] == VKL_FORMAT_TILE ? 1 : 512;
      uint64_t expected     = voxels;
      if (anyTemporal) {
        Data *ti = tuIdxD ? tuIdxD->item(i) : nullptr;
        Data *tt = tuTmD ? tuTmD->item(i) : nullptr;
        expected = verifyTemporalData(*device, voxels, tformat[i], steps[i], ti, tt);
        dataOff[i] = dataTotal;
        dataTotal += expected;
        if (tformat[i] == VKL_TEMPORAL_FORMAT_STRUCTURED) leafSteps[i] = (uint32_t)steps[i];
        if (tformat[i] == VKL_TEMPORAL_FORMAT_UNSTRUCTURED) {
          idxOff[i]   = flatIdx.size();
          timesOff[i] = flatTimes.size();
          if (ti->dataType == VKL_UINT) {
            const std::vector<uint32_t> v = ti->hostArray<uint32_t>();
            flatIdx.insert(flatIdx.end(), v.begin(), v.end());
          } else {
            const std::vector<uint64_t> v = ti->hostArray<uint64_t>();
            flatIdx.insert(flatIdx.end(), v.begin(), v.end());
          }
          const std::vector<float> tv = tt->hostArray<float>();
          flatTimes.insert(flatTimes.end(), tv.begin(), tv.end());
        }
      }
      for (uint32_t a = 0; a < na; a++) {
        Data *ad = multi ? ld->item(a) : ld;
        if (ad->numItems < expected) {
          std::ostringstream ss;
          ss << "Node data too small: found " << ad->numItems << " elements, but expected " << expected;
          throw VklError(ss.str());
        }
        if (ad->numItems > expected) {
          std::ostringstream ss;
          ss << "Node data too big: found " << ad->numItems << " elements, but expected " << expected;
          throw VklError(ss.str());
        }
        if (ad->dataType != types[a]) {
          std::ostringstream ss;
          ss << "inconsistent leaf attribute data type (expected " << (int)types[a] << ")";
          throw VklError(ss.str());
        }
        std::vector<uint8_t> h = ad->hostCopy();
        auto &dst = (format[i] == VKL_FORMAT_TILE && !anyTemporal) ? tile[a] : dense[a];
        dst.insert(dst.end(), h.begin(), h.end());
      }
    }
    grid.temporalFormat = anyTemporal ? 1u : 0u;
    if (anyTemporal) {
      auto up = [&](const void *src, size_t bytes) {
        DeviceBuffer *b = newBuffer(std::max<size_t>(bytes, 8));
        if (bytes) b->upload(src, bytes, st);
        return b->ptr;
      };
      grid.leafDataOffset  = static_cast<const uint64_t *>(up(dataOff.data(), n * 8));
      grid.leafSteps       = static_cast<const uint32_t *>(up(leafSteps.data(), n * 4));
      grid.leafIdxOffset   = static_cast<const uint64_t *>(up(idxOff.data(), n * 8));
      grid.leafTimesOffset = static_cast<const uint64_t *>(up(timesOff.data(), n * 8));
      grid.tuIndices       = up(flatIdx.data(), flatIdx.size() * 8);
      grid.tuIndex64       = 1;
      grid.tuTimes         = static_cast<const float *>(up(flatTimes.data(), flatTimes.size() * 4));
    }
    for (uint32_t a = 0; a < na; a++) {
      AttrView av{};
      av.type = (uint32_t)types[a];
      DeviceBuffer *bd = newBuffer(dense[a].size());
      bd->upload(dense[a].data(), dense[a].size(), st);
      av.data = bd->ptr;
      if (!tile[a].empty()) {
        DeviceBuffer *bt = newBuffer(tile[a].size());
        bt->upload(tile[a].data(), tile[a].size(), st);
        av.tile = bt->ptr;
      }
      attrViews.push_back(av);
    }
    DeviceBuffer *bs = newBuffer(n * sizeof(uint64_t));
    bs->upload(slot.data(), n * sizeof(uint64_t), st);
    grid.leafSlot = bs->as<uint64_t>();
  }
  for (VKLDataType t : types)
    if (t != VKL_HALF && t != VKL_FLOAT) {
      std::ostringstream ss;
      ss << "node.data arrays have data type " << (int)t << " but only " << (int)VKL_HALF << " (VKL_HALF), "
         << (int)VKL_FLOAT << " (VKL_FLOAT), or " << (int)VKL_DATA << " (VKL_DATA) is supported for vdb volumes.";
      throw VklError(ss.str());
    }
  grid.numAttributes = (uint32_t)attrViews.size();
  const std::vector<float> bg = parseBackground(*this, grid.numAttributes);
  for (uint32_t a = 0; a < grid.numAttributes; a++) attrViews[a].background = bg[a];
  uploadAttrViews(*this, st);

  float col[12] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f};
  getAffine3f("indexToObject", col);
  buildTransforms(col, i2o, grid.o2i);

  // computeBbox / computeRootOrigin (VdbVolume.cpp:71-80, 567-581).
  int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
  for (uint64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) {
      lo[c] = std::min<int64_t>(lo[c], origin[3 * i + c]);
      hi[c] = std::max<int64_t>(hi[c], (int64_t)origin[3 * i + c] + kRes[level[i]]);
    }
  for (int c = 0; c < 3; c++)
    if (hi[c] - lo[c] > (int64_t)kRes[0]) throw VklError("input leaves do not fit into a single root level node");
  for (int c = 0; c < 3; c++) {
    grid.rootOrigin[c] = (int32_t)kRes[1] * (int32_t)std::floor((int32_t)lo[c] / (float)kRes[1]);
    grid.activeSize[c] = (uint32_t)(hi[c] - grid.rootOrigin[c]);
  }
  float ibox[6] = {(float)lo[0], (float)lo[1], (float)lo[2], (float)hi[0], (float)hi[1], (float)hi[2]};
  int32_t clip[6];
  if (getBox3i("indexClippingBounds", clip) && clip[0] <= clip[3] && clip[1] <= clip[4] && clip[2] <= clip[5])
    for (int c = 0; c < 6; c++) ibox[c] = (float)clip[c];
  for (int c = 0; c < 3; c++) {
    grid.domainBox[c]     = ibox[c] - (float)grid.rootOrigin[c];
    grid.domainBox[3 + c] = ibox[3 + c] - (float)grid.rootOrigin[c];
  }
  transformBounds(i2o, ibox, bbox);
  grid.dense            = 0;
  grid.constantCellData = 1;
  grid.maxSamplingDepth = (uint32_t)maxSamplingDepth;

  std::vector<uint32_t> offs(3 * n);
  for (uint64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) offs[3 * i + c] = (uint32_t)(origin[3 * i + c] - grid.rootOrigin[c]);
  denseLeafOrigins.clear();
  for (uint64_t i = 0; i < n; i++)
    if (format[i] == VKL_FORMAT_DENSE_ZYX) denseLeafOrigins.insert(denseLeafOrigins.end(), &origin[3 * i], &origin[3 * i] + 3);
  // Leaf codes carry the packed index for packed input, else the input index.
  std::vector<uint64_t> leafId(n);
  for (uint64_t i = 0; i < n; i++) leafId[i] = packedD ? slot[i] : i;
  // Voxel codes carry the data format and the temporal format.
  std::vector<uint32_t> codeFormat(n);
  for (uint64_t i = 0; i < n; i++) codeFormat[i] = (format[i] & 3u) | ((tformat[i] & 3u) << 2);
  finishGrid(*this, level, codeFormat, offs, leafId, st);

  // "cornerPacked" (bool): a leaf index grid over the domain plus, per dense
  // leaf, 512 cell records of the 8 corner values the sampler reads (8 floats
  // per cell; baked for this volume's maxSamplingDepth).
  //
  // "leafIndexGrid" (int; -1 automatic, 0 off, 1 on): the same leaf index
  // grid without the records, which full-depth sampling uses as a flattened
  // node table (one lookup per stencil instead of three dependent node
  // loads). Automatic: built when it takes at most 1 GiB (4 bytes per 8^3
  // cell of the domain: 512 MiB for a 4096^3 domain).
  grid.leafIndex   = nullptr;
  grid.packedDepth = 0xffffffffu;
  const bool cornerPacked = getBool("cornerPacked", false);
  if (cornerPacked && anyTemporal)
    throw VklError("cornerPacked is not supported for temporal data", VKL_INVALID_ARGUMENT);
  for (int c = 0; c < 3; c++) grid.leafDims[c] = (grid.activeSize[c] + 7) / 8;
  const uint64_t cellsOfGrid = (uint64_t)grid.leafDims[0] * grid.leafDims[1] * grid.leafDims[2];
  const int indexMode        = getInt("leafIndexGrid", -1);
  const bool indexGrid = !anyTemporal && (cornerPacked || indexMode > 0 ||
                                          (indexMode < 0 && cellsOfGrid * sizeof(uint32_t) <= (1ull << 30)));
  if (indexGrid) {
    const uint64_t numDense = denseLeafOrigins.size() / 3;
    DeviceBuffer origins(std::max<size_t>(denseLeafOrigins.size() * sizeof(int32_t), 4));
    if (numDense) origins.upload(denseLeafOrigins.data(), denseLeafOrigins.size() * sizeof(int32_t), st);
    // tile nodes per input level 1..3
    std::vector<int32_t> tileOrg[3];
    for (uint64_t i = 0; i < n; i++)
      if (format[i] == VKL_FORMAT_TILE && level[i] >= 1 && level[i] <= 3)
        tileOrg[level[i] - 1].insert(tileOrg[level[i] - 1].end(), &origin[3 * i], &origin[3 * i] + 3);
    DeviceBuffer tileBuf[3];
    const int32_t *tilePtr[3] = {nullptr, nullptr, nullptr};
    uint64_t numTiles[3]      = {0, 0, 0};
    for (int l = 0; l < 3; l++) {
      numTiles[l] = tileOrg[l].size() / 3;
      if (!numTiles[l]) continue;
      tileBuf[l].allocate(tileOrg[l].size() * sizeof(int32_t));
      tileBuf[l].upload(tileOrg[l].data(), tileOrg[l].size() * sizeof(int32_t), st);
      tilePtr[l] = tileBuf[l].as<int32_t>();
    }
    DeviceBuffer *index = newBuffer(cellsOfGrid * sizeof(uint32_t));
    checkHip(launchBuildLeafIndex(grid, origins.as<int32_t>(), numDense, tilePtr, numTiles, index->as<uint32_t>(), st),
             "leaf index");
    checkHip(hipStreamSynchronize(st), "leaf index");
    grid.leafIndex = index->as<uint32_t>();
  }
  if (cornerPacked) {
    const uint64_t numDense = denseLeafOrigins.size() / 3;
    DeviceBuffer origins(denseLeafOrigins.size() * sizeof(int32_t));
    origins.upload(denseLeafOrigins.data(), denseLeafOrigins.size() * sizeof(int32_t), st);
    for (uint32_t a = 0; a < grid.numAttributes; a++) {
      DeviceBuffer *cells = newBuffer(numDense * 512 * 8 * sizeof(float));
      checkHip(launchPackLeafCells(grid, attrViews[a], a, origins.as<int32_t>(), numDense, cells->as<float>(), st),
               "cell records");
      attrViews[a].cells = cells->as<float>();
    }
    checkHip(hipStreamSynchronize(st), "corner packing");
    uploadAttrViews(*this, st);
    grid.packedDepth = grid.maxSamplingDepth;
  }
  committed = true;
}

}  // namespace vklhip
