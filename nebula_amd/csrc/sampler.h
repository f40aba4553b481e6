// sampler.h -- library-internal access to a learned sampler's device views.
#pragma once

#include "nbl_host.h"
#include "nbl_device.h"
#include "mlp_device.h"

namespace nbl {
// The two networks (index NBL_SAMPLER_ELASTIC / NBL_SAMPLER_INELASTIC), or NULL.
const dmlp* sampler_views(nbl_sampler s);
// The material set the sampler was trained for.
nbl_materials sampler_materials(nbl_sampler s);
// Bit k set: network k replaces table k in the transport.
uint32_t sampler_kinds(nbl_sampler s);
}
