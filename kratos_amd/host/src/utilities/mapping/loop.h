// Reference include path of the host API; the declarations live in
// kratos/types.h.
#pragma once
#include "../../../kratos/types.h"
