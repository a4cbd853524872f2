// Shared internals of the C ABI: error reporting and the opaque handles.
#pragma once

#include <string>
#include <vector>

#include "../../include/mdgpu_c.h"
#include "symbolic.h"

namespace mdgpu {

// Set the calling thread's last error and return `code`.
int fail(int code, const std::string& msg);
void clear_error();

}  // namespace mdgpu

struct mdgpu_symbolic {
    mdgpu::Symbolic S;
    std::vector<int> level;      // height of each front (leaves 0)
    int nlevels = 0;
    int max_front = 0;
};
