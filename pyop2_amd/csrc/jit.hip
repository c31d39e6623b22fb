// Runtime compilation and launch of generated gather/compute/scatter
// kernels (user C-string kernels that have no hand-written implementation).
#include <hip/hiprtc.h>

#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace op2hip {

struct Module {
  hipModule_t mod = nullptr;
  std::mutex lock;
  std::unordered_map<std::string, hipFunction_t> fns;
};

static int compile_code(const char *source, const char *const *options, int32_t noptions,
                        std::vector<char> *code, char *log, size_t log_bytes) {
  OP2HIP_REQUIRE(source, "jit: NULL source");
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, source, "op2_wrapper.hip", 0, nullptr, nullptr);
  if (r != HIPRTC_SUCCESS) {
    set_error("hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
    return OP2HIP_ERR_JIT;
  }
  std::vector<const char *> opts;
  bool has_arch = false;
  for (int i = 0; i < noptions; ++i) {
    opts.push_back(options[i]);
    if (std::strncmp(options[i], "--offload-arch", 14) == 0) has_arch = true;
  }
  if (!has_arch) opts.push_back("--offload-arch=gfx950");
  r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  size_t n = 0;
  hiprtcGetProgramLogSize(prog, &n);
  if (log && log_bytes) {
    log[0] = 0;
    if (n > 1) {
      std::vector<char> buf(n + 1, 0);
      hiprtcGetProgramLog(prog, buf.data());
      std::strncpy(log, buf.data(), log_bytes - 1);
      log[log_bytes - 1] = 0;
    }
  }
  if (r != HIPRTC_SUCCESS) {
    set_error("hiprtcCompileProgram: %s", hiprtcGetErrorString(r));
    hiprtcDestroyProgram(&prog);
    return OP2HIP_ERR_JIT;
  }
  size_t sz = 0;
  hiprtcGetCodeSize(prog, &sz);
  code->resize(sz);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return OP2HIP_OK;
}

}  // namespace op2hip

using namespace op2hip;

extern "C" {

int op2hip_jit_code_object(const char *source, const char *const *options, int32_t noptions,
                           void *code, size_t *code_bytes, char *log, size_t log_bytes) {
  OP2HIP_REQUIRE(code_bytes, "jit_code_object: NULL size");
  std::vector<char> obj;
  int rc = compile_code(source, options, noptions, &obj, log, log_bytes);
  if (rc) return rc;
  if (!code) {
    *code_bytes = obj.size();
    return OP2HIP_OK;
  }
  if (*code_bytes < obj.size()) {
    set_error("jit_code_object: buffer %zu < %zu bytes", *code_bytes, obj.size());
    *code_bytes = obj.size();
    return OP2HIP_ERR_WORKSPACE;
  }
  std::memcpy(code, obj.data(), obj.size());
  *code_bytes = obj.size();
  return OP2HIP_OK;
}

int op2hip_jit_load(const void *code, void **module) {
  OP2HIP_REQUIRE(code && module, "jit_load: NULL argument");
  auto *m = new Module();
  hipError_t e = hipModuleLoadData(&m->mod, code);
  if (e != hipSuccess) {
    delete m;
    set_error("hipModuleLoadData: %s", hipGetErrorString(e));
    return OP2HIP_ERR_JIT;
  }
  *module = m;
  return OP2HIP_OK;
}

int op2hip_jit_compile(const char *source, const char *const *options, int32_t noptions,
                       void **module, char *log, size_t log_bytes) {
  std::vector<char> obj;
  int rc = compile_code(source, options, noptions, &obj, log, log_bytes);
  if (rc) return rc;
  return op2hip_jit_load(obj.data(), module);
}

int op2hip_jit_unload(void *module) {
  if (!module) return OP2HIP_OK;
  auto *m = static_cast<Module *>(module);
  hipError_t e = m->mod ? hipModuleUnload(m->mod) : hipSuccess;
  delete m;
  if (e != hipSuccess) {
    set_error("hipModuleUnload: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  return OP2HIP_OK;
}

int op2hip_jit_launch(void *module, const char *name, uint32_t grid, uint32_t block,
                      uint32_t shared_bytes, void **args, void *stream) {
  OP2HIP_REQUIRE(module && name, "jit_launch: NULL argument");
  auto *m = static_cast<Module *>(module);
  hipFunction_t fn;
  {
    std::lock_guard<std::mutex> g(m->lock);
    auto it = m->fns.find(name);
    if (it == m->fns.end()) {
      hipError_t e = hipModuleGetFunction(&fn, m->mod, name);
      if (e != hipSuccess) {
        set_error("hipModuleGetFunction(%s): %s", name, hipGetErrorString(e));
        return OP2HIP_ERR_JIT;
      }
      m->fns.emplace(name, fn);
    } else {
      fn = it->second;
    }
  }
  if (grid == 0) return OP2HIP_OK;
  OP2HIP_CHECK(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, shared_bytes,
                                     stream_of(stream), args, nullptr));
  return OP2HIP_OK;
}

}  // extern "C"
