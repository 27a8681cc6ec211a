// Streaming triplet reader (files -> pinned ring -> device u/v/r arrays).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace cumf {

// The files read_problem would read for `path` (shards path0..path79 when any
// exist, else path itself) and their whole-record counts (size / 12).
int triplet_files(const char* path, std::vector<std::string>* files, std::vector<int64_t>* counts, bool* sharded);

// Streams the files' records, in file order, into the device arrays u, v, r
// (room for the sum of counts) on stream st.  mm_out receives the largest u,
// the largest v and the smallest id (-1, -1, INT_MAX when there are none).
int stream_triplets(const std::vector<std::string>& files, const std::vector<int64_t>& counts, int32_t* u, int32_t* v,
                    float* r, hipStream_t st, int32_t mm_out[3]);

}  // namespace cumf
