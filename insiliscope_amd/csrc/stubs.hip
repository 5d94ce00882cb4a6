// stubs.hip -- teardown hooks for subsystems not built into this library yet.
#include "iscg_internal.h"

namespace iscg {

void destroy_roi(iscg_context* h) { h->roi = nullptr; }
}  // namespace iscg
