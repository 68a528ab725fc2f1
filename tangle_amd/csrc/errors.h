// Error reporting shared by the host translation units of the library.
#pragma once

#include <string>

namespace tgh {

// Records `message` as the calling thread's last error (tgh_last_error) and
// returns `code`.
int fail(int code, const std::string &message);

}  // namespace tgh
