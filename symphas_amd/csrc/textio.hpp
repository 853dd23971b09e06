// textio.hpp - the reference's data files (GNU checkpoint text, GNU plot
// matrix, CSV, XDR) of one device field, shared by the single-field and the
// coupled-field plans.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/sps_fd.h"

namespace sps {

// One field on the device: its values (unpadded, x fastest), the geometry and
// dtype in d (dim, n, h, dtype are used), its size and the stream to work on.
struct FieldView {
  void* data;
  sps_fd_desc d;
  size_t nsites;
  hipStream_t stream;
};

int file_checkpoint_text(const FieldView& v, int index, const double* origin,
                         char* text, size_t capacity, size_t* length);
int file_plot_text(const FieldView& v, int format, const double* origin,
                   char* text, size_t capacity, size_t* length);
int file_checkpoint_read(const FieldView& v, const char* text, size_t length,
                         int* index);
int file_xdr_write(const FieldView& v, int index, const double* origin,
                   void* data, size_t capacity, size_t* length);
int file_xdr_read(const FieldView& v, const void* data, size_t length,
                  int* index);

}  // namespace sps
