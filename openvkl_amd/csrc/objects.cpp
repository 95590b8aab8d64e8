// Devices, typed parameters and data arrays of the HIP device.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <sstream>

#include "objects.h"

namespace vklhip {

void checkHip(hipError_t e, const char *what)
{
  if (e == hipSuccess) return;
  std::ostringstream ss;
  ss << "HIP error in " << what << ": " << hipGetErrorString(e);
  throw VklError(ss.str(), e == hipErrorOutOfMemory ? VKL_OUT_OF_MEMORY : VKL_UNKNOWN_ERROR);
}

// Element sizes of VKLDataType (openvkl/common/VKLCommon.cpp:20-150); managed
// object handles are stored as VKLObject {host, device}.
size_t sizeOfType(VKLDataType t)
{
  switch (t) {
  case VKL_DEVICE:
  case VKL_VOID_PTR:
  case VKL_STRING: return sizeof(void *);
  case VKL_BOOL: return sizeof(bool);
  case VKL_OBJECT:
  case VKL_DATA:
  case VKL_VOLUME: return sizeof(VKLObject);
  case VKL_CHAR:
  case VKL_UCHAR: return 1;
  case VKL_VEC2C:
  case VKL_VEC2UC: return 2;
  case VKL_VEC3C:
  case VKL_VEC3UC: return 3;
  case VKL_VEC4C:
  case VKL_VEC4UC: return 4;
  case VKL_SHORT:
  case VKL_USHORT:
  case VKL_HALF: return 2;
  case VKL_VEC2S:
  case VKL_VEC2US:
  case VKL_VEC2H: return 4;
  case VKL_VEC3S:
  case VKL_VEC3US:
  case VKL_VEC3H: return 6;
  case VKL_VEC4S:
  case VKL_VEC4US:
  case VKL_VEC4H: return 8;
  case VKL_INT:
  case VKL_UINT:
  case VKL_FLOAT: return 4;
  case VKL_VEC2I:
  case VKL_VEC2UI:
  case VKL_VEC2F: return 8;
  case VKL_VEC3I:
  case VKL_VEC3UI:
  case VKL_VEC3F: return 12;
  case VKL_VEC4I:
  case VKL_VEC4UI:
  case VKL_VEC4F: return 16;
  case VKL_LONG:
  case VKL_ULONG:
  case VKL_DOUBLE: return 8;
  case VKL_VEC2L:
  case VKL_VEC2UL:
  case VKL_VEC2D: return 16;
  case VKL_VEC3L:
  case VKL_VEC3UL:
  case VKL_VEC3D: return 24;
  case VKL_VEC4L:
  case VKL_VEC4UL:
  case VKL_VEC4D: return 32;
  case VKL_BOX1I:
  case VKL_BOX1F: return 8;
  case VKL_BOX2I:
  case VKL_BOX2F: return 16;
  case VKL_BOX3I:
  case VKL_BOX3F: return 24;
  case VKL_BOX4I:
  case VKL_BOX4F: return 32;
  case VKL_LINEAR2F: return 16;
  case VKL_LINEAR3F: return 36;
  case VKL_AFFINE2F: return 24;
  case VKL_AFFINE3F: return 48;
  default: break;
  }
  std::ostringstream ss;
  ss << "cannot compute size of unknown VKLDataType " << (int)t;
  throw VklError(ss.str());
}

bool isDevicePointer(const void *p)
{
  if (!p) return false;
  hipPointerAttribute_t attr;
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory: clear the sticky error
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

bool isPinnedHost(const void *p)
{
  if (!p) return false;
  hipPointerAttribute_t attr;
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // pageable host memory: clear the sticky error
    return false;
  }
  return attr.type == hipMemoryTypeHost;
}

// ---- DeviceBuffer ------------------------------------------------------------

void DeviceBuffer::allocate(size_t n)
{
  reset();
  if (n == 0) return;
  checkHip(hipMalloc(&ptr, n), "hipMalloc");
  bytes = n;
}

void DeviceBuffer::reset()
{
  if (ptr) (void)hipFree(ptr);
  ptr   = nullptr;
  bytes = 0;
}

void DeviceBuffer::upload(const void *src, size_t n, hipStream_t s)
{
  if (n == 0) return;
  checkHip(hipMemcpyAsync(ptr, src, n, hipMemcpyDefault, s), "hipMemcpyAsync");
  checkHip(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// ---- ManagedObject -------------------------------------------------------------

ManagedObject::~ManagedObject()
{
  for (auto &kv : params)
    if (kv.second.object) kv.second.object->release();
}

void ManagedObject::setParam(const std::string &name, VKLDataType type, const void *mem)
{
  if (!mem) throw VklError("null parameter value for '" + name + "'");
  Param p;
  p.type = type;
  if (type == VKL_DATA || type == VKL_VOLUME || type == VKL_OBJECT) {
    const VKLObject *h = static_cast<const VKLObject *>(mem);
    p.object           = static_cast<ManagedObject *>(h->host);
    if (p.object) p.object->retain();
  } else if (type == VKL_STRING) {
    p.str = static_cast<const char *>(mem);
  } else {
    const size_t n = sizeOfType(type);
    p.bytes.assign(static_cast<const uint8_t *>(mem), static_cast<const uint8_t *>(mem) + n);
  }
  auto it = params.find(name);
  if (it != params.end() && it->second.object) it->second.object->release();
  params[name] = std::move(p);
}

const ManagedObject::Param *ManagedObject::findParam(const std::string &name) const
{
  auto it = params.find(name);
  return it == params.end() ? nullptr : &it->second;
}

int ManagedObject::getInt(const std::string &name, int def) const
{
  const Param *p = findParam(name);
  if (!p) return def;
  if (p->type == VKL_INT || p->type == VKL_UINT) {
    int v;
    std::memcpy(&v, p->bytes.data(), 4);
    return v;
  }
  if (p->type == VKL_BOOL) return p->bytes[0] ? 1 : 0;
  return def;
}

float ManagedObject::getFloat(const std::string &name, float def) const
{
  const Param *p = findParam(name);
  if (!p || p->type != VKL_FLOAT) return def;
  float v;
  std::memcpy(&v, p->bytes.data(), 4);
  return v;
}

bool ManagedObject::getBool(const std::string &name, bool def) const
{
  const Param *p = findParam(name);
  if (!p) return def;
  if (p->type == VKL_BOOL) return p->bytes[0] != 0;
  if (p->type == VKL_INT || p->type == VKL_UINT) return getInt(name, 0) != 0;
  return def;
}

bool ManagedObject::getVec3i(const std::string &name, int32_t out[3]) const
{
  const Param *p = findParam(name);
  if (!p || (p->type != VKL_VEC3I && p->type != VKL_VEC3UI)) return false;
  std::memcpy(out, p->bytes.data(), 12);
  return true;
}

bool ManagedObject::getVec3f(const std::string &name, float out[3]) const
{
  const Param *p = findParam(name);
  if (!p || p->type != VKL_VEC3F) return false;
  std::memcpy(out, p->bytes.data(), 12);
  return true;
}

bool ManagedObject::getBox3i(const std::string &name, int32_t out[6]) const
{
  const Param *p = findParam(name);
  if (!p || p->type != VKL_BOX3I) return false;
  std::memcpy(out, p->bytes.data(), 24);
  return true;
}

// "indexToObject": a VKL_AFFINE3F value or a float data array of >= 12
// entries, both column-major (vx, vy, vz, p) as in
// openvkl/devices/cpu/volume/vdb/VdbVolume.h:35-57.
bool ManagedObject::getAffine3f(const std::string &name, float out[12]) const
{
  const Param *p = findParam(name);
  if (!p) return false;
  if (p->type == VKL_AFFINE3F) {
    std::memcpy(out, p->bytes.data(), 48);
    return true;
  }
  if (p->type == VKL_DATA && p->object) {
    const Data *d = static_cast<const Data *>(p->object);
    if (d->dataType == VKL_FLOAT && d->numItems >= 12) {
      std::vector<float> v = d->hostArray<float>();
      std::memcpy(out, v.data(), 48);
      return true;
    }
  }
  return false;
}

const void *ManagedObject::getVoidPtr(const std::string &name) const
{
  const Param *p = findParam(name);
  if (!p || p->type != VKL_VOID_PTR) return nullptr;
  const void *v;
  std::memcpy(&v, p->bytes.data(), sizeof(v));
  return v;
}

Data *ManagedObject::getData(const std::string &name) const
{
  const Param *p = findParam(name);
  if (!p || p->type != VKL_DATA) return nullptr;
  return static_cast<Data *>(p->object);
}

// ---- Data -------------------------------------------------------------------------

Data::Data(Device *d, size_t n, VKLDataType t, const void *src, VKLDataCreationFlags flags, size_t byteStride)
    : ManagedObject(d), numItems(n), dataType(t)
{
  if (n == 0) throw VklError("numItems must be positive");
  if (!src) throw VklError("source cannot be null");
  const bool managed = (t == VKL_DATA || t == VKL_VOLUME || t == VKL_OBJECT);
  const size_t es    = sizeOfType(t);
  stride             = byteStride ? byteStride : es;
  if (managed) {
    for (size_t i = 0; i < n; i++) {
      const VKLObject *h = reinterpret_cast<const VKLObject *>(static_cast<const uint8_t *>(src) + i * stride);
      ManagedObject *o   = static_cast<ManagedObject *>(h->host);
      if (o) o->retain();
      objects.push_back(o);
    }
    stride = es;
    return;
  }
  onDevice = isDevicePointer(src);
  if (flags & VKL_DATA_SHARED_BUFFER) {
    base = static_cast<const uint8_t *>(src);
    return;
  }
  // Copy, compacting strides.
  if (onDevice) {
    if (stride != es) throw VklError("strided device-resident data must be created with VKL_DATA_SHARED_BUFFER");
    owned.resize(n * es);
    checkHip(hipMemcpy(owned.data(), src, n * es, hipMemcpyDeviceToHost), "hipMemcpy");
    onDevice = false;
  } else {
    owned.resize(n * es);
    const uint8_t *s = static_cast<const uint8_t *>(src);
    if (stride == es)
      std::memcpy(owned.data(), s, n * es);
    else
      for (size_t i = 0; i < n; i++) std::memcpy(owned.data() + i * es, s + i * stride, es);
  }
  stride = es;
  base   = owned.data();
}

Data::~Data()
{
  for (ManagedObject *o : objects)
    if (o) o->release();
}

std::vector<uint8_t> Data::hostCopy() const
{
  const size_t es = sizeOfType(dataType);
  std::vector<uint8_t> out(numItems * es);
  if (onDevice) {
    if (stride == es)
      checkHip(hipMemcpy(out.data(), base, out.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    else
      checkHip(hipMemcpy2D(out.data(), es, base, stride, es, numItems, hipMemcpyDeviceToHost), "hipMemcpy2D");
    return out;
  }
  if (stride == es)
    std::memcpy(out.data(), base, out.size());
  else
    for (size_t i = 0; i < numItems; i++) std::memcpy(out.data() + i * es, base + i * stride, es);
  return out;
}

// ---- Device -----------------------------------------------------------------------

namespace {
void defaultErrorCallback(void *, VKLError e, const char *msg)
{
  const char *out = std::getenv("OPENVKL_ERROR_OUTPUT");
  if (out && std::string(out) == "cout")
    std::fprintf(stdout, "OPENVKL ERROR (%d): %s\n", (int)e, msg);
  else if (!out || std::string(out) != "none")
    std::fprintf(stderr, "OPENVKL ERROR (%d): %s\n", (int)e, msg);
}
}  // namespace

Device::Device() : errorCallback(defaultErrorCallback) {}

Device::~Device()
{
  if (pinned) (void)hipHostFree(pinned);
}

// Device params (openvkl/api/Device.cpp:135-209): logLevel, deviceIndex
// (HIP ordinal) and stream (hipStream_t; default = the null stream).
void Device::commit()
{
  auto geti = [&](const char *n, int def) {
    auto it = params.find(n);
    if (it == params.end() || it->second.bytes.size() < 4) return def;
    int v;
    std::memcpy(&v, it->second.bytes.data(), 4);
    return v;
  };
  logLevel = geti("logLevel", VKL_LOG_WARNING);
  if (const char *e = std::getenv("OPENVKL_LOG_LEVEL")) logLevel = std::atoi(e);
  hipDevice = geti("deviceIndex", 0);
  auto it   = params.find("stream");
  if (it != params.end() && it->second.bytes.size() == sizeof(void *))
    std::memcpy(&stream, it->second.bytes.data(), sizeof(void *));
  int count = 0;
  checkHip(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (hipDevice < 0 || hipDevice >= count) throw VklError("invalid deviceIndex for the hip device");
  makeCurrent();
  hipDeviceProp_t prop;
  checkHip(hipGetDeviceProperties(&prop, hipDevice), "hipGetDeviceProperties");
  std::ostringstream ss;
  ss << "hip device " << hipDevice << ": " << prop.gcnArchName << ", " << prop.multiProcessorCount << " CUs";
  log(VKL_LOG_DEBUG, ss.str());
  committed = true;
}

void Device::makeCurrent() const { (void)hipSetDevice(hipDevice); }

void Device::handleError(VKLError e, const std::string &msg)
{
  lastErrorCode    = e;
  lastErrorMessage = msg;
  if (errorCallback) errorCallback(errorUserData, e, msg.c_str());
}

void Device::log(int level, const std::string &msg)
{
  if (level < logLevel) return;
  if (logCallback)
    logCallback(logUserData, msg.c_str());
  else
    std::fprintf(stderr, "[openvkl hip] %s\n", msg.c_str());
}

}  // namespace vklhip
