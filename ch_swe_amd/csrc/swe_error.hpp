// Error reporting shared by libchswe's translation units: records the
// message swe_last_error() returns on this thread and passes the code on.
#pragma once

#include <string>

namespace swe {
int set_error(int code, const std::string &msg);
}
