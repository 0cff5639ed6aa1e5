// host_common.h -- shared by the host-side sources of steps_host.h.
#pragma once

#include <cstddef>
#include <string>

#include "steps_force.h"
#include "steps_host.h"

// Records the message for steps_host_last_error() and returns code.
int host_fail(int code, const std::string &msg);
