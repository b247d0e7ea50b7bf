// probgen.h -- problem-generator entry points (user/probgen.h of the
// reference): a generator defines prob::run; prob::cmd reads every
// command-line parameter file into args, later files overriding earlier
// keys (user/probgen.cpp:8-16).
#pragma once

#include "input.h"

namespace prob {
void cmd(int argc, char *argv[], input &args);
void run(int argc, char *argv[]);
}  // namespace prob
