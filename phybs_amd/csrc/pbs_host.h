// pbs_host.h -- host-side helpers shared by the C-ABI translation units.
#pragma once
#include <string>

namespace pbs {
// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string &msg);
}  // namespace pbs
