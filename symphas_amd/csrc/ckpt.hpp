// ckpt.hpp - the checkpoint directory of a run (host code): file names, the
// progress record and resuming from a recorded index.
#pragma once

#include <functional>
#include <string>

#include "common.hpp"

namespace sps {

// Formats field `id`'s checkpoint text (GNU checkpoint file: full header line
// then the values) for solution index `index`.
using CkptFormat = std::function<int(int index, std::string& text)>;
// Parses one GNU checkpoint text (full header line then values) into the
// field; *index receives the header's index.
using CkptParse = std::function<int(const char* text, size_t length, int* index)>;

// Write field `id` at `index` under dir/checkpoint (single_file: one file
// per field holding every checkpoint, else one file per field and index).
int ckpt_write_field(const char* dir, int id, int index, int single_file,
                     const CkptFormat& format);
// Append the index (and, when done, the completion key) to the progress
// record dir/checkpoint/progress.txt.
int ckpt_record(const char* dir, int index, int done);
// Read field `id`'s checkpoint for `index` (single_file: the last block whose
// index does not pass `index`) through parse; *index_read gets its index.
int ckpt_read_field(const char* dir, int id, int index, int single_file,
                    int* index_read, const CkptParse& parse);

}  // namespace sps
