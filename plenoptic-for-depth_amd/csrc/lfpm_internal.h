// lfpm_internal.h — declarations shared by the host and device translation units.
#pragma once
#include "../../include/lfpm.h"

namespace lfpm {
int set_error(int code, const char* fmt, ...);
const char* last_error();
int resolve_crop(const lfpm_layout_params* p);
}  // namespace lfpm
