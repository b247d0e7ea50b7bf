// Reference include path of the host API; the declarations live in
// kratos/probgen.h.
#pragma once
#include "../../kratos/probgen.h"
