// osy_internal.hpp - shared declarations of the host library.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/opensayal.h"

namespace osy {

// Records `msg` as the calling thread's last error and returns `status`.
int set_error(int status, const std::string& msg);
void clear_error();

void config_defaults(osy_config* c, int32_t width, int32_t height);
int config_parse_text(const std::string& text, osy_config* out);
int config_parse_file(const char* path, osy_config* out);

// Checks a configuration against what the solver implements.  Returns
// OSY_OK or an error status with the message set.
int config_validate(const osy_config& c);

}  // namespace osy
