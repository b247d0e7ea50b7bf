// Reference include path of the host API; the declarations live in
// kratos/mesh.h.
#pragma once
#include "../../kratos/mesh.h"
