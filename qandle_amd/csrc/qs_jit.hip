// Run-time specialised pass kernels: hipRTC compilation and launch.
//
// The host planner emits the HIP source of one kernel per pass with the pass
// program as compile-time constants (tile layout, register maps, table
// indices; CNOTs between register bits become register renames).  These entry
// points compile such a source for the device's architecture, load the code
// object, look up kernels and launch them with a by-value argument block.
// Code objects can be cached by the caller (qs_jit_compile returns the bytes,
// qs_jit_load takes them back).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <stdlib.h>
#include <string.h>

#include "qs_common.h"

namespace {

struct JitCode {
  char *code = nullptr;
  size_t size = 0;
  char *log = nullptr;
  size_t log_size = 0;
};

int save_hip(hipError_t e) {
  if (e == hipSuccess) return QS_OK;
  qs::g_last_hip_error = (int)e;
  return QS_EHIP;
}

}  // namespace

extern "C" {

int qs_jit_compile(const char *src, const char *name, const char *const *options, int noptions, void **handle) {
  if (!src || !name || !handle || noptions < 0 || (noptions > 0 && !options)) return QS_EINVAL;
  *handle = nullptr;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src, name, 0, nullptr, nullptr) != HIPRTC_SUCCESS) return QS_EHIP;
  const hiprtcResult rc = hiprtcCompileProgram(prog, noptions, options);
  JitCode *jc = new JitCode();
  size_t ls = 0;
  if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 0) {
    jc->log = (char *)malloc(ls + 1);
    if (jc->log && hiprtcGetProgramLog(prog, jc->log) == HIPRTC_SUCCESS) {
      jc->log[ls] = 0;
      jc->log_size = ls;
    }
  }
  if (rc == HIPRTC_SUCCESS) {
    size_t cs = 0;
    if (hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0) {
      jc->code = (char *)malloc(cs);
      if (jc->code && hiprtcGetCode(prog, jc->code) == HIPRTC_SUCCESS) jc->size = cs;
    }
  }
  hiprtcDestroyProgram(&prog);
  *handle = jc;
  return jc->size ? QS_OK : QS_EHIP;
}

int64_t qs_jit_code_size(const void *handle) { return handle ? (int64_t)((const JitCode *)handle)->size : -1; }

int qs_jit_code_copy(const void *handle, void *out, int64_t size) {
  const JitCode *jc = (const JitCode *)handle;
  if (!jc || !out || size < (int64_t)jc->size) return QS_EINVAL;
  memcpy(out, jc->code, jc->size);
  return QS_OK;
}

int64_t qs_jit_log_size(const void *handle) { return handle ? (int64_t)((const JitCode *)handle)->log_size : -1; }

int qs_jit_log_copy(const void *handle, char *out, int64_t size) {
  const JitCode *jc = (const JitCode *)handle;
  if (!jc || !out || size < 1) return QS_EINVAL;
  const size_t n = jc->log_size < (size_t)(size - 1) ? jc->log_size : (size_t)(size - 1);
  if (n) memcpy(out, jc->log, n);
  out[n] = 0;
  return QS_OK;
}

void qs_jit_release_code(void *handle) {
  JitCode *jc = (JitCode *)handle;
  if (!jc) return;
  free(jc->code);
  free(jc->log);
  delete jc;
}

int qs_jit_load(const void *code, int64_t size, void **module) {
  if (!code || size <= 0 || !module) return QS_EINVAL;
  hipModule_t m;
  const int rc = save_hip(hipModuleLoadData(&m, code));
  *module = rc == QS_OK ? (void *)m : nullptr;
  return rc;
}

int qs_jit_function(void *module, const char *name, void **function) {
  if (!module || !name || !function) return QS_EINVAL;
  hipFunction_t f;
  const int rc = save_hip(hipModuleGetFunction(&f, (hipModule_t)module, name));
  *function = rc == QS_OK ? (void *)f : nullptr;
  return rc;
}

int64_t qs_jit_resident_blocks(void *function, int block) {
  if (!function || block < 1) return -1;
  int per_cu = 0, cus = 0, dev = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (hipFunction_t)function, block, 0) != hipSuccess)
    return -1;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  return (int64_t)(per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
}

int qs_jit_launch(void *function, unsigned grid, unsigned block, const void *args, int64_t args_size,
                  void *stream) {
  if (!function || grid < 1 || block < 1 || !args || args_size <= 0) return QS_EINVAL;
  size_t sz = (size_t)args_size;
  void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void *>(args), HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz,
                    HIP_LAUNCH_PARAM_END};
  return save_hip(hipModuleLaunchKernel((hipFunction_t)function, grid, 1, 1, block, 1, 1, 0, (hipStream_t)stream,
                                        nullptr, config));
}

int qs_jit_unload(void *module) {
  if (!module) return QS_EINVAL;
  return save_hip(hipModuleUnload((hipModule_t)module));
}

}  // extern "C"
