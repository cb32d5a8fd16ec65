// Library runtime: per-thread error state, the library stream, and the
// device workspace cache.
#include "prg_internal.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>

namespace prg {

namespace {
thread_local int t_code = OK;
thread_local std::string t_msg;

struct Runtime {
  std::mutex mu;
  bool init = false;
  int init_status = OK;
  hipStream_t stream = nullptr;
  std::map<std::string, std::pair<void*, size_t>> ws;
  unsigned generation = 1;
  void* host_stage = nullptr;
  size_t host_stage_bytes = 0;
};

Runtime& rt() {
  static Runtime r;
  return r;
}
}  // namespace

void set_error(int code, const std::string& msg) {
  t_code = code;
  t_msg = msg;
}
void clear_error() {
  t_code = OK;
  t_msg.clear();
}
int last_error_code() { return t_code; }
const char* last_error_message() { return t_msg.c_str(); }

int ensure_device() {
  Runtime& r = rt();
  std::lock_guard<std::mutex> lk(r.mu);
  if (r.init) {
    if (r.init_status != OK) set_error(r.init_status, "GPU runtime initialisation failed");
    return r.init_status;
  }
  r.init = true;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    r.init_status = ERR_NO_DEVICE;
    set_error(ERR_NO_DEVICE, "no HIP device visible (this library has no CPU fallback)");
    return r.init_status;
  }
  e = hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    r.init_status = ERR_HIP;
    set_error(ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    return r.init_status;
  }
  return OK;
}

bool device_ready() {
  Runtime& r = rt();
  std::lock_guard<std::mutex> lk(r.mu);
  return r.init && r.init_status == OK;
}

hipStream_t stream() { return rt().stream; }

void* workspace(const char* name, size_t bytes, int* status) {
  Runtime& r = rt();
  std::lock_guard<std::mutex> lk(r.mu);
  auto it = r.ws.find(name);
  if (it != r.ws.end() && it->second.second >= bytes) return it->second.first;
  if (it != r.ws.end()) {
    // the stream may still use the old buffer
    hipStreamSynchronize(r.stream);
    hipFree(it->second.first);
    r.ws.erase(it);
    ++r.generation;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
  if (e != hipSuccess) {
    set_error(ERR_HIP, std::string("hipMalloc workspace ") + name + ": " + hipGetErrorString(e));
    if (status) *status = ERR_HIP;
    return nullptr;
  }
  hipMemsetAsync(p, 0, bytes ? bytes : 256, r.stream);
  r.ws[name] = {p, bytes};
  return p;
}

unsigned workspace_generation() {
  Runtime& r = rt();
  std::lock_guard<std::mutex> lk(r.mu);
  return r.generation;
}

void* host_staging(size_t bytes, int* status) {
  Runtime& r = rt();
  std::lock_guard<std::mutex> lk(r.mu);
  if (r.host_stage && r.host_stage_bytes >= bytes) return r.host_stage;
  if (r.host_stage) {
    hipStreamSynchronize(r.stream);
    hipHostFree(r.host_stage);
    r.host_stage = nullptr;
  }
  hipError_t e = hipHostMalloc(&r.host_stage, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    set_error(ERR_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    if (status) *status = ERR_HIP;
    return nullptr;
  }
  r.host_stage_bytes = bytes;
  return r.host_stage;
}

void release_workspaces() {
  Runtime& r = rt();
  std::lock_guard<std::mutex> lk(r.mu);
  if (!r.init || r.init_status != OK) return;
  hipStreamSynchronize(r.stream);
  for (auto& kv : r.ws) hipFree(kv.second.first);
  r.ws.clear();
  ++r.generation;
  if (r.host_stage) hipHostFree(r.host_stage);
  r.host_stage = nullptr;
  r.host_stage_bytes = 0;
}

}  // namespace prg

extern "C" {

int prg_last_error_code(void) { return prg::last_error_code(); }
const char* prg_last_error(void) { return prg::last_error_message(); }

void prg_version(void) { std::printf("progress-mi355x density-matrix solvers (gfx950)\n"); }

void prg_progress_init(void) {
  prg::clear_error();
  if (prg::ensure_device() == prg::OK) prg::noalloc_setup();
}

void prg_progress_shutdown(void) {
  prg::clear_error();
  prg::release_workspaces();
}

}  // extern "C"
