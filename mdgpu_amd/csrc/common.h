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
    // independent matrices of a batch (one segment for a single matrix):
    // columns [seg_col[k], seg_col[k+1]) and caller entries
    // [seg_nnz[k], seg_nnz[k+1]) belong to matrix k
    std::vector<int> seg_col, seg_nnz;
};
