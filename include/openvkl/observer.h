/* Observers: data returned from volumes and samplers to the application;
 * declarations match openvkl/include/openvkl/observer.h:21-75.
 *
 * Volume observers ("vdb" volumes): "InnerNode" (int parameter "maxDepth",
 * default 1) returns one (6 + 2 * numAttributes)-float tuple per node at or
 * above maxDepth: object-space bounding box, then the value range of every
 * attribute (VdbInnerNodeObserver.cpp:96-186).
 * Sampler observers (samplers on "vdb" volumes): "LeafNodeAccess" returns one
 * uint32 per input node; entry i is 1 when a traversal of this sampler ended
 * on leaf index i after the observer was created (VdbLeafAccessObserver.cpp,
 * VdbSampleInner.ih.in:69-74). The mapped array is a snapshot taken by
 * vklMapObserver. */
#pragma once

#include "common.h"
#include "sampler.h"
#include "volume.h"

#ifdef __cplusplus
struct VKLObserver : public VKLObject
{
};
extern "C" {
#else
typedef VKLObject VKLObserver;
#endif

NOWARN_C_LINKAGE_PUSH

/* observer.h:37 */
OPENVKL_INTERFACE VKLObserver vklNewVolumeObserver(VKLVolume volume, const char *type);

/* observer.h:40 */
OPENVKL_INTERFACE VKLObserver vklNewSamplerObserver(VKLSampler sampler, const char *type);

NOWARN_C_LINKAGE_POP

/* observer.h:51 */
OPENVKL_INTERFACE const void *vklMapObserver(VKLObserver observer);

/* observer.h:57 */
OPENVKL_INTERFACE void vklUnmapObserver(VKLObserver observer);

/* observer.h:63 */
OPENVKL_INTERFACE VKLDataType vklGetObserverElementType(VKLObserver observer);

/* observer.h:69 */
OPENVKL_INTERFACE size_t vklGetObserverElementSize(VKLObserver observer);

/* observer.h:75 */
OPENVKL_INTERFACE size_t vklGetObserverNumElements(VKLObserver observer);

#ifdef __cplusplus
}
#endif
