// Reference include path of the host API; the declarations live in
// kratos/dynamics.h.
#pragma once
#include "../../../../kratos/dynamics.h"
