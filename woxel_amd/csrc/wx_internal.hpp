// Shared internals of libwoxel_hip: error reporting and opaque handle types.
#pragma once

#include <string>

#include "../../include/woxel_hip.h"
#include "vdb_host.hpp"

struct wx_vdb {
    wx::HostVdb tree;
};

namespace wx {

// Records the message for wx_last_error() on this thread and returns code.
int set_error(int code, const std::string &msg);
int ok();

// Frees the batch scratch kept for `stream` (a hipStream_t about to be destroyed).
void drop_stream_scratch(void *stream);

}  // namespace wx
