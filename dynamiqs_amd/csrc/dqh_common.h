// Shared host-side helpers of the dqh C-ABI: thread-local error message.
#pragma once
#include <string>

namespace dqh {
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
}  // namespace dqh
