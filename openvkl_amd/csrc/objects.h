// Object model of the HIP device: devices, typed parameters, data arrays,
// volumes, samplers and iterator contexts. Mirrors the reference's
// ManagedObject / Device split (openvkl/common/ManagedObject.h:20-196,
// openvkl/api/Device.h:42-185) with HIP device memory behind each object.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "openvkl/device/openvkl.h"
#include "openvkl/openvkl.h"
#include "vkl_views.h"

namespace vklhip {

// An exception carrying a VKLError code (default VKL_UNKNOWN_ERROR, as the
// reference maps every std::exception, openvkl/api/API.cpp:65-89).
struct VklError : std::runtime_error
{
  VKLError code;
  explicit VklError(const std::string &m, VKLError c = VKL_UNKNOWN_ERROR) : std::runtime_error(m), code(c) {}
};

void checkHip(hipError_t e, const char *what);
// Object-space bounds of an index-space box: the 8 corners through i2o
// (row-major 3x4), component-wise min / max.
void transformBounds(const float i2o[12], const float box[6], float out[6]);
size_t sizeOfType(VKLDataType t);
bool isDevicePointer(const void *p);
// Page-locked host memory (hipHostMalloc'd or registered): DMA-able in place.
bool isPinnedHost(const void *p);

// Owning device allocation.
struct DeviceBuffer
{
  void *ptr    = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { allocate(n); }
  ~DeviceBuffer() { reset(); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  void allocate(size_t n);
  void reset();
  void upload(const void *src, size_t n, hipStream_t s);
  template <typename T>
  T *as() const { return static_cast<T *>(ptr); }
};

struct Device;

struct ManagedObject
{
  struct Param
  {
    VKLDataType type = VKL_UNKNOWN;
    std::vector<uint8_t> bytes;
    ManagedObject *object = nullptr;  // retained, for VKL_DATA / VKL_VOLUME
    std::string str;
  };

  std::atomic<int> refs{1};
  Device *device = nullptr;
  std::map<std::string, Param> params;

  explicit ManagedObject(Device *d) : device(d) {}
  virtual ~ManagedObject();
  virtual void commit() {}
  virtual std::string toString() const = 0;

  void retain() { refs++; }
  void release()
  {
    if (--refs == 0) delete this;
  }

  void setParam(const std::string &name, VKLDataType type, const void *mem);
  bool hasParam(const std::string &name) const { return params.count(name) != 0; }
  const Param *findParam(const std::string &name) const;

  int getInt(const std::string &name, int def) const;
  float getFloat(const std::string &name, float def) const;
  bool getBool(const std::string &name, bool def) const;
  bool getVec3i(const std::string &name, int32_t out[3]) const;
  bool getVec3f(const std::string &name, float out[3]) const;
  bool getBox3i(const std::string &name, int32_t out[6]) const;
  bool getAffine3f(const std::string &name, float out[12]) const;  // column-major vx, vy, vz, p
  const void *getVoidPtr(const std::string &name) const;
  struct Data *getData(const std::string &name) const;
};

// vklNewData: copied (strides compacted) unless VKL_DATA_SHARED_BUFFER.
// Shared buffers may live in device memory.
struct Data : ManagedObject
{
  size_t numItems      = 0;
  VKLDataType dataType = VKL_UNKNOWN;
  size_t stride        = 0;  // bytes between elements
  bool onDevice        = false;
  const uint8_t *base  = nullptr;
  std::vector<uint8_t> owned;
  std::vector<ManagedObject *> objects;  // VKL_DATA / VKL_VOLUME elements (retained)

  Data(Device *d, size_t n, VKLDataType t, const void *src, VKLDataCreationFlags flags, size_t byteStride);
  ~Data() override;
  std::string toString() const override { return "openvkl::Data"; }
  bool compact() const { return stride == sizeOfType(dataType); }
  size_t byteSize() const { return numItems * sizeOfType(dataType); }
  // Compact host copy of the elements (downloads device-resident buffers).
  std::vector<uint8_t> hostCopy() const;
  template <typename T>
  std::vector<T> hostArray() const
  {
    std::vector<uint8_t> b = hostCopy();
    std::vector<T> r(numItems);
    if (!b.empty()) std::memcpy(r.data(), b.data(), numItems * sizeof(T));
    return r;
  }
  Data *item(size_t i) const { return static_cast<Data *>(objects.at(i)); }
};

struct Volume : ManagedObject
{
  using ManagedObject::ManagedObject;

  GridView grid{};
  float i2o[12]{};                  // index -> object, row-major 3x4
  std::vector<int32_t> denseLeafOrigins;  // vdb: index-space origins of dense leaves, packed-slot order
  std::vector<AttrView> attrViews;  // host copies of grid.attrs
  float bbox[6]{};
  std::vector<Range1f> valueRanges;
  int filter            = kFilterLinear;
  int gradientFilter    = kFilterLinear;
  int maxSamplingDepth  = 3;
  bool committed        = false;

  // Device memory owned by the committed volume.
  std::vector<std::unique_ptr<DeviceBuffer>> buffers;
  // Host mirror of the node table (observers, tests).
  std::vector<uint64_t> hostVoxels[3];
  std::vector<Range1f> hostRanges[3];
  uint64_t numLeaves = 0;

  DeviceBuffer *newBuffer(size_t bytes);
  uint32_t numAttributes() const { return grid.numAttributes; }
  void commitCommonParams();
};

struct DenseVolume : Volume
{
  using Volume::Volume;
  void commit() override;
  std::string toString() const override { return "openvkl::DenseVdbVolume"; }
};

struct VdbVolume : Volume
{
  using Volume::Volume;
  void commit() override;
  std::string toString() const override { return "openvkl::VdbVolume"; }
};

struct Sampler : ManagedObject
{
  Volume *volume;
  SamplerView view{};
  DeviceBuffer deviceView;  // SamplerView in device memory (the VKLSampler.device handle)
  bool committed = false;
  // Spatially binned stream sampling ("binnedSampling": -1 automatic, 0 off,
  // 1 whenever the volume and filter allow it). One scratch per sampler,
  // serialised by binMutex (calls are synchronous).
  int binnedMode = -1;
  std::mutex binMutex;
  DeviceBuffer binBuf;
  // LeafNodeAccess observation: one uint32 per leaf index holding the stamp
  // of the last traversal that ended on that leaf. accessEpoch counts the
  // observers ever created (each observer keeps its own epoch and reads an
  // entry as accessed when its stamp is at least that epoch); traversals
  // store the current epoch while any observer is alive. The buffer lives as
  // long as the sampler, so views copied into plans stay valid.
  DeviceBuffer accessBuf;
  uint32_t accessEpoch = 0;
  uint32_t liveObservers = 0;
  uint64_t viewVersion   = 0;  // bumped whenever view changes
  // Guards view / viewVersion against observers created or released while
  // other threads iterate (contexts copy view under it).
  std::mutex viewMutex;

  Sampler(Device *d, Volume *v);
  ~Sampler() override;
  void commit() override;
  std::string toString() const override { return "openvkl::Sampler"; }
  const AttrView &attr(uint32_t a) const;
  // A copy of view taken under viewMutex (what one call launches with).
  SamplerView viewCopy()
  {
    std::lock_guard<std::mutex> lock(viewMutex);
    return view;
  }
  // Applies the observation state to view (observed samplers traverse the
  // node table: corner-packed cell records bypass it) and uploads deviceView.
  void applyObservation();
};

struct IteratorContext : ManagedObject
{
  Sampler *sampler;
  IteratorContextView view{};
  DeviceBuffer rangesBuf, valuesBuf;
  DeviceBuffer deviceView;  // copy of `view` in device memory (hit kernels)
  uint64_t samplerVersion = 0;  // Sampler::viewVersion that view.sampler copies
  bool isHit;
  bool committed = false;

  IteratorContext(Device *d, Sampler *s, bool hit);
  ~IteratorContext() override;
  void commit() override;
  // Picks up sampler state changed since commit (re-commit, observers) and
  // returns a copy of the view, both under viewMutex: host threads iterating
  // one context while an observer is created or released each launch with a
  // consistent view.
  IteratorContextView refreshSampler();
  std::mutex viewMutex;
  // Work counter of the persistent batch kernels (4 bytes, allocated on the
  // first batch call); batchMutex serialises batch calls on this context.
  DeviceBuffer batchWork;
  std::mutex batchMutex;
  std::string toString() const override { return isHit ? "openvkl::HitIteratorContext" : "openvkl::IntervalIteratorContext"; }
};

// vklNew*Observer (openvkl/include/openvkl/observer.h:21-75). Observers
// retain their target.
struct Observer : ManagedObject
{
  using ManagedObject::ManagedObject;
  virtual const void *map()           = 0;
  virtual void unmap() {}
  virtual VKLDataType elementType() const = 0;
  virtual size_t elementSize() const      = 0;
  virtual size_t numElements() const      = 0;
  std::string toString() const override { return "openvkl::Observer"; }
};

// "InnerNode" on vdb volumes (VdbInnerNodeObserver.cpp:96-186): built on
// commit, or on the first map after creation.
struct InnerNodeObserver : Observer
{
  Volume *volume;
  std::vector<float> nodes;
  uint32_t numFloats = 0;
  bool built         = false;

  InnerNodeObserver(Device *d, Volume *v);
  ~InnerNodeObserver() override;
  void commit() override;
  const void *map() override;
  VKLDataType elementType() const override { return VKL_OBJECT; }
  size_t elementSize() const override { return sizeof(float) * numFloats; }
  size_t numElements() const override { return numFloats ? nodes.size() / numFloats : 0; }
};

// "LeafNodeAccess" on samplers of vdb volumes (VdbLeafAccessObserver.cpp).
struct LeafAccessObserver : Observer
{
  Sampler *sampler;
  uint32_t epoch;
  std::vector<uint32_t> snapshot;

  LeafAccessObserver(Device *d, Sampler *s);
  ~LeafAccessObserver() override;
  const void *map() override;
  VKLDataType elementType() const override { return VKL_UINT; }
  size_t elementSize() const override { return sizeof(uint32_t); }
  size_t numElements() const override { return snapshot.size(); }
};

struct Device
{
  std::atomic<int> refs{1};
  std::map<std::string, ManagedObject::Param> params;
  bool committed  = false;
  int hipDevice   = 0;
  hipStream_t stream = nullptr;
  int logLevel    = VKL_LOG_WARNING;

  VKLError lastErrorCode = VKL_NO_ERROR;
  std::string lastErrorMessage;
  VKLErrorCallback errorCallback = nullptr;
  void *errorUserData            = nullptr;
  VKLLogCallback logCallback     = nullptr;
  void *logUserData              = nullptr;

  // Page-locked host staging of host-pointer calls on pageable arrays (two
  // chunk slots), allocated on first use and kept until the device is
  // released; stagingMutex serialises such calls per device.
  std::mutex stagingMutex;
  void *pinned       = nullptr;
  size_t pinnedBytes = 0;

  Device();
  ~Device();
  void commit();
  void handleError(VKLError e, const std::string &msg);
  void log(int level, const std::string &msg);
  void makeCurrent() const;
};

}  // namespace vklhip
