/* Umbrella header: same role as openvkl/include/openvkl/openvkl.h. */
#pragma once

#include "VKLBackgroundUndefined.h"
#include "VKLDataType.h"
#include "VKLError.h"
#include "VKLFilter.h"
#include "VKLFormat.h"
#include "VKLLogLevel.h"
#include "VKLTemporalFormat.h"
#include "common.h"
#include "data.h"
#include "device.h"
#include "iterator.h"
#include "observer.h"
#include "parameters.h"
#include "sampler.h"
#include "vdb.h"
#include "volume.h"

#define OPENVKL_VERSION_MAJOR 2
#define OPENVKL_VERSION_MINOR 0
#define OPENVKL_VERSION_PATCH 2
