// Library-wide C-ABI entry points: version and last-error message.
#include <string>

#include "hp_common.hpp"

namespace hp {
std::string& last_error_slot() {
  static thread_local std::string msg;
  return msg;
}
}  // namespace hp

extern "C" {

int hp_abi_version(void) { return HP_ABI_VERSION; }

const char* hp_last_error(void) { return hp::last_error_slot().c_str(); }

void hp_struct_sizes(int64_t* out) {
  if (!out) return;
  out[0] = (int64_t)sizeof(hp_ctrl);
  out[1] = (int64_t)sizeof(hp_grid_desc);
  out[2] = (int64_t)sizeof(hp_ensemble);
}

}  // extern "C"
