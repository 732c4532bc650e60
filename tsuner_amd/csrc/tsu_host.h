// tsu_host.h -- host helpers shared by the runtime files (internal).
#pragma once

#include <string>

#include "../../include/tsuner.h"
#include "tsu_device.h"

namespace tsu {

// Run constants from the reference parameters (tsu_api.cpp).
void make_dev_params(const tsu_sim_params* p, const tsu_scheme* s, DevParams* P);
// Validates a scheme; on failure records the message and returns the code.
int validate_scheme(const tsu_scheme* s);
// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string& msg);

}  // namespace tsu
