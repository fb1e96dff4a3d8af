// Runtime part of the op2hip C-ABI: device selection, hiprtc compilation of
// generated loop sources with an on-disk code-object cache, module loading,
// kernel launch and constant upload.
//
// Replaces the JIT step of the reference's backends (instant + gcc in
// pyop2/host.py:225-261, pycuda + nvcc in pyop2/cuda.py:602-651): the loop
// source is generated in Python and compiled here for gfx950.
#include <hip/hiprtc.h>

#include <cerrno>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "common.h"

namespace op2hip {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[4096];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

// FNV-1a 64-bit over the source, options and architecture: the key of the
// on-disk code-object cache.
static uint64_t fnv1a(const void* data, size_t n, uint64_t h) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

static bool read_file(const std::string& path, std::vector<char>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.seekg(0, std::ios::end);
  std::streamoff n = f.tellg();
  if (n <= 0) return false;
  out.resize(static_cast<size_t>(n));
  f.seekg(0);
  f.read(out.data(), n);
  return static_cast<bool>(f);
}

static void write_file_atomic(const std::string& path, const std::vector<char>& data) {
  std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    f.write(data.data(), static_cast<std::streamsize>(data.size()));
    if (!f) return;
  }
  rename(tmp.c_str(), path.c_str());
}

static void mkdirs(const std::string& dir) {
  std::string cur;
  std::stringstream ss(dir);
  std::string part;
  if (!dir.empty() && dir[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    mkdir(cur.c_str(), 0755);
  }
}

struct Module {
  hipModule_t mod = nullptr;
  std::vector<char> code;
};

}  // namespace op2hip

using namespace op2hip;

extern "C" {

const char* op2hip_last_error(void) { return g_last_error.c_str(); }

int op2hip_init(int device) {
  int n = 0;
  OP2HIP_CHECK(hipGetDeviceCount(&n));
  OP2HIP_REQUIRE(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
  OP2HIP_CHECK(hipSetDevice(device));
  OP2HIP_CHECK(hipFree(nullptr));
  return OP2HIP_OK;
}

// Limits used to size launches: LDS bytes per workgroup, compute units,
// LDS bytes per compute unit.
int op2hip_device_limits(int device, int* lds_per_block, int* num_cu, int* lds_per_cu) {
  int v = 0;
  if (lds_per_block) {
    OP2HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    *lds_per_block = v;
  }
  if (num_cu) {
    OP2HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device));
    *num_cu = v;
  }
  if (lds_per_cu) {
    OP2HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                       device));
    *lds_per_cu = v;
  }
  return OP2HIP_OK;
}

int op2hip_device_arch(int device, char* out, int outlen) {
  hipDeviceProp_t prop;
  OP2HIP_CHECK(hipGetDeviceProperties(&prop, device));
  snprintf(out, static_cast<size_t>(outlen), "%s", prop.gcnArchName);
  return OP2HIP_OK;
}

// Compile `src` with hiprtc (or load it from `cache_dir`) and return the
// code object in a malloc'd buffer owned by the caller (free with
// op2hip_free_host).  Needs no GPU, so generated sources can be checked on
// a build host.
int op2hip_compile(const char* src, const char* const* opts, int nopts, const char* cache_dir,
                   void** code_out, size_t* code_size, int* cache_hit) {
  OP2HIP_REQUIRE(src && code_out && code_size, "null argument to op2hip_compile");
  uint64_t h = 1469598103934665603ull;
  h = fnv1a(src, strlen(src), h);
  for (int i = 0; i < nopts; ++i) h = fnv1a(opts[i], strlen(opts[i]) + 1, h);
  int ver = 0;
  hiprtcVersion(&ver, &ver);
  h = fnv1a(&ver, sizeof(ver), h);
  char key[32];
  snprintf(key, sizeof(key), "%016llx", static_cast<unsigned long long>(h));
  std::string path;
  std::vector<char> code;
  if (cache_hit) *cache_hit = 0;
  if (cache_dir && cache_dir[0]) {
    path = std::string(cache_dir) + "/" + key + ".co";
    if (read_file(path, code)) {
      if (cache_hit) *cache_hit = 1;
    }
  }
  if (code.empty()) {
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, src, "op2_loop.hip", 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) {
      set_error("hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
      return OP2HIP_ERR_COMPILE;
    }
    r = hiprtcCompileProgram(prog, nopts, opts);
    if (r != HIPRTC_SUCCESS) {
      size_t ls = 0;
      hiprtcGetProgramLogSize(prog, &ls);
      std::string log(ls + 1, '\0');
      hiprtcGetProgramLog(prog, &log[0]);
      hiprtcDestroyProgram(&prog);
      set_error("hiprtc compilation failed: %s\n%s", hiprtcGetErrorString(r), log.c_str());
      return OP2HIP_ERR_COMPILE;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    if (!path.empty()) {
      mkdirs(cache_dir);
      write_file_atomic(path, code);
    }
  }
  void* buf = malloc(code.size());
  if (!buf) {
    set_error("out of host memory");
    return OP2HIP_ERR_ALLOC;
  }
  memcpy(buf, code.data(), code.size());
  *code_out = buf;
  *code_size = code.size();
  return OP2HIP_OK;
}

void op2hip_free_host(void* p) { free(p); }

int op2hip_module_load(const void* code, size_t size, void** mod_out) {
  OP2HIP_REQUIRE(code && size && mod_out, "null argument to op2hip_module_load");
  Module* m = new Module();
  m->code.assign(static_cast<const char*>(code), static_cast<const char*>(code) + size);
  hipError_t e = hipModuleLoadData(&m->mod, m->code.data());
  if (e != hipSuccess) {
    delete m;
    set_error("hipModuleLoadData: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  *mod_out = m;
  return OP2HIP_OK;
}

int op2hip_module_destroy(void* mod) {
  Module* m = static_cast<Module*>(mod);
  if (!m) return OP2HIP_OK;
  if (m->mod) hipModuleUnload(m->mod);
  delete m;
  return OP2HIP_OK;
}

int op2hip_module_function(void* mod, const char* name, void** fn_out) {
  Module* m = static_cast<Module*>(mod);
  OP2HIP_REQUIRE(m && name && fn_out, "null argument to op2hip_module_function");
  hipFunction_t f;
  OP2HIP_CHECK(hipModuleGetFunction(&f, m->mod, name));
  *fn_out = reinterpret_cast<void*>(f);
  return OP2HIP_OK;
}

int op2hip_module_global(void* mod, const char* name, void** dptr, size_t* bytes) {
  Module* m = static_cast<Module*>(mod);
  OP2HIP_REQUIRE(m && name && dptr && bytes, "null argument to op2hip_module_global");
  hipDeviceptr_t p;
  size_t n = 0;
  OP2HIP_CHECK(hipModuleGetGlobal(&p, &n, m->mod, name));
  *dptr = reinterpret_cast<void*>(p);
  *bytes = n;
  return OP2HIP_OK;
}

// Copy host bytes to a module global (a kernel Const) on `stream` and wait:
// the host buffer may be reused as soon as this returns.
int op2hip_set_global(void* dptr, const void* src, size_t bytes, void* stream) {
  OP2HIP_REQUIRE(dptr && src, "null argument to op2hip_set_global");
  hipStream_t s = static_cast<hipStream_t>(stream);
  OP2HIP_CHECK(hipMemcpyHtoDAsync(reinterpret_cast<hipDeviceptr_t>(dptr), const_cast<void*>(src),
                                  bytes, s));
  OP2HIP_CHECK(hipStreamSynchronize(s));
  return OP2HIP_OK;
}

namespace {
constexpr int kSmallBytes = 256;
struct SmallBytes {
  unsigned char b[kSmallBytes];
};

// Lane i stores byte i: the bytes travel as the kernel's argument.
__global__ void __launch_bounds__(kSmallBytes) store_small_bytes(unsigned char* __restrict__ dst,
                                                                  SmallBytes v, int n) {
  const int i = static_cast<int>(threadIdx.x);
  if (i < n) dst[i] = v.b[i];
}
}  // namespace

// Store up to 256 host bytes at `dptr` in `stream` order, without waiting:
// the bytes are copied into a kernel argument before this returns, so the
// host buffer may be reused at once and the store may be captured in a
// graph.  A Global's value reaches the device this way.
int op2hip_store_small(void* dptr, const void* src, size_t bytes, void* stream) {
  OP2HIP_REQUIRE(dptr && src, "null argument to op2hip_store_small");
  OP2HIP_REQUIRE(bytes <= static_cast<size_t>(kSmallBytes),
                 "op2hip_store_small takes at most %d bytes (got %zu)", kSmallBytes, bytes);
  if (bytes == 0) return OP2HIP_OK;
  SmallBytes v;
  std::memset(v.b, 0, sizeof(v.b));
  std::memcpy(v.b, src, bytes);
  hipLaunchKernelGGL(store_small_bytes, dim3(1), dim3(kSmallBytes), 0,
                     static_cast<hipStream_t>(stream), static_cast<unsigned char*>(dptr), v,
                     static_cast<int>(bytes));
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

// Compile, load and look up in one call: `src` through hiprtc (with the
// code-object cache), its module loaded, `entry` resolved.  The module is
// the caller's (op2hip_module_destroy).
int op2hip_module_compile(const char* src, const char* entry, const char* const* opts, int nopts,
                          const char* cache_dir, void** mod_out, void** fn_out) {
  OP2HIP_REQUIRE(src && entry && mod_out && fn_out, "null argument to op2hip_module_compile");
  void* code = nullptr;
  size_t size = 0;
  int hit = 0;
  int st = op2hip_compile(src, opts, nopts, cache_dir, &code, &size, &hit);
  if (st != OP2HIP_OK) return st;
  st = op2hip_module_load(code, size, mod_out);
  op2hip_free_host(code);
  if (st != OP2HIP_OK) return st;
  st = op2hip_module_function(*mod_out, entry, fn_out);
  if (st != OP2HIP_OK) {
    op2hip_module_destroy(*mod_out);
    *mod_out = nullptr;
  }
  return st;
}

// Set a module global (a kernel Const) by name from `bytes` host bytes.
int op2hip_module_set_const(void* mod, const char* name, const void* host, size_t bytes,
                            void* stream) {
  void* dptr = nullptr;
  size_t size = 0;
  int st = op2hip_module_global(mod, name, &dptr, &size);
  if (st != OP2HIP_OK) return st;
  OP2HIP_REQUIRE(bytes <= size, "Const %s holds %zu bytes, %zu given", name, size, bytes);
  return op2hip_set_global(dptr, host, bytes, stream);
}

// Launch a generated loop kernel whose parameters are all 8 bytes wide
// (pointers and 64-bit integers): argv[i] is the i-th parameter's value.
int op2hip_launch(void* fn, unsigned grid, unsigned block, unsigned lds, void* stream,
                  const uint64_t* argv, int nargs) {
  OP2HIP_REQUIRE(fn, "null kernel");
  OP2HIP_REQUIRE(nargs >= 0 && nargs <= 256, "too many kernel arguments (%d)", nargs);
  if (grid == 0) return OP2HIP_OK;
  void* params[256];
  for (int i = 0; i < nargs; ++i) params[i] = const_cast<uint64_t*>(argv + i);
  OP2HIP_CHECK(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(fn), grid, 1, 1, block, 1, 1,
                                     lds, static_cast<hipStream_t>(stream), params, nullptr));
  return OP2HIP_OK;
}

// Kernel attributes for the launch heuristics: registers and spills.
int op2hip_function_attrs(void* fn, int* num_regs, int* local_bytes, int* max_threads) {
  hipFunction_t f = reinterpret_cast<hipFunction_t>(fn);
  OP2HIP_REQUIRE(f, "null kernel");
  int v = 0;
  if (num_regs) {
    OP2HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, f));
    *num_regs = v;
  }
  if (local_bytes) {
    OP2HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f));
    *local_bytes = v;
  }
  if (max_threads) {
    OP2HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, f));
    *max_threads = v;
  }
  return OP2HIP_OK;
}

// Stream-ordered timing of the work enqueued between two calls: the events
// are recorded on the same stream the loop kernels are launched on.
int op2hip_event_create(void** ev) {
  hipEvent_t e;
  OP2HIP_CHECK(hipEventCreate(&e));
  *ev = e;
  return OP2HIP_OK;
}

int op2hip_event_destroy(void* ev) {
  if (ev) hipEventDestroy(static_cast<hipEvent_t>(ev));
  return OP2HIP_OK;
}

int op2hip_event_record(void* ev, void* stream) {
  OP2HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(ev), static_cast<hipStream_t>(stream)));
  return OP2HIP_OK;
}

int op2hip_event_elapsed_ms(void* start, void* stop, float* ms) {
  OP2HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  OP2HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                   static_cast<hipEvent_t>(stop)));
  return OP2HIP_OK;
}

int op2hip_stream_sync(void* stream) {
  OP2HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return OP2HIP_OK;
}

}  // extern "C"
