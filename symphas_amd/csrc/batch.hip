// batch.hip - batched API entry points (filled in by the persistent kernel).
#include "common.hpp"
#include "plan.hpp"

extern "C" {

int sps_fd_batch_create(const sps_fd_desc*, int, sps_fd_batch** out) {
  if (out) *out = nullptr;
  sps::set_error("batched stepping not built yet");
  return SPS_E_UNSUPPORTED;
}
int sps_fd_batch_destroy(sps_fd_batch*) { return SPS_OK; }
int sps_fd_batch_upload(sps_fd_batch*, int, const void*, int) {
  sps::set_error("batched stepping not built yet");
  return SPS_E_UNSUPPORTED;
}
int sps_fd_batch_download(sps_fd_batch*, int, void*, int) {
  sps::set_error("batched stepping not built yet");
  return SPS_E_UNSUPPORTED;
}
int sps_fd_batch_step(sps_fd_batch*, int) {
  sps::set_error("batched stepping not built yet");
  return SPS_E_UNSUPPORTED;
}
int sps_fd_batch_reduce(sps_fd_batch*, double*, double*) {
  sps::set_error("batched stepping not built yet");
  return SPS_E_UNSUPPORTED;
}
int sps_fd_batch_sync(sps_fd_batch*) {
  sps::set_error("batched stepping not built yet");
  return SPS_E_UNSUPPORTED;
}

}  // extern "C"
