// Snapshot I/O (restart files) of device or host vectors; see snapshot.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/exadg_amd.h"

namespace ExaDG
{
void rotate_snapshots(std::string const & path, int n_keep);
void write_snapshot(std::string const & path, exadg_snapshot_info const & info,
                    void const * const * vectors, bool on_device, int n_keep, hipStream_t s);
exadg_snapshot_info read_snapshot_info(std::string const & path);
exadg_snapshot_info read_snapshot(std::string const & path, void * const * vectors, int n_vectors,
                                  int64_t n, int precision, bool on_device, hipStream_t s);
}  // namespace ExaDG
