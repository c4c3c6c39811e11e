// device_render.h -- shading data of the wavefront renderer in HBM.
#pragma once

#include <stdint.h>

#include "device_scene.h"

namespace az {

struct RenderState {
    bool ready = false;
};

}  // namespace az
