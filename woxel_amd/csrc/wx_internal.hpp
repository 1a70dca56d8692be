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

// Checks a layout descriptor (every index a kernel derives from it stays inside its
// pool), allocates the grid's pools on the current device and copies the root table,
// node records and, when d->leaves is set, the leaf words from the host. With
// d->leaves NULL the leaf pool (n3 x 256 bytes, grid_pool(g, 3)) is the caller's to
// fill. `who` prefixes error messages.
int grid_new(const wx_grid_desc *d, const char *who, wx_grid **out);
// Device address of pool k of a grid (0 root, 1 node-5, 2 node-4, 3 leaves).
void *grid_pool(wx_grid *g, int k);

}  // namespace wx
