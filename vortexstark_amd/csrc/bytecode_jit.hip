// Compiled tier of the bytecode constraint evaluator: the program becomes
// straight-line HIP source (bc_jit_source: every value a local variable,
// trace loads read once per row, constants and row-offset steps literal),
// compiled for gfx950 with hipRTC on first use and cached per device,
// program and domain shape. The compiler keeps the values in VGPRs, so a row
// costs the arithmetic of its program, with no instruction decode and no
// register-file traffic.
#include <hip/hiprtc.h>

#include <mutex>
#include <string>
#include <unordered_map>

#include "vx_bytecode.h"
#include "vx_internal.h"

namespace vx {

namespace {

const char kFieldHeader[] =
#include "vx_field_text.inc"
    ;

struct JitKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

std::mutex g_mu;
std::unordered_map<std::string, JitKernel> g_kernels;
bool g_enabled = true;

void rtc_check(hiprtcResult r, const char* what) {
  if (r != HIPRTC_SUCCESS) throw Error(ERR_INTERNAL, std::string("bytecode JIT: ") + what + ": " + hiprtcGetErrorString(r));
}

// hipRTC compile of the generated source to a gfx950 code object.
std::string compile_code(const std::string& src) {
  // the field header without its runtime includes (hipRTC supplies them)
  std::string hdr = kFieldHeader;
  for (const char* inc : {"#include <hip/hip_runtime.h>", "#include <stdint.h>"}) {
    const size_t p = hdr.find(inc);
    if (p != std::string::npos) hdr.erase(p, std::char_traits<char>::length(inc));
  }
  const char* hs[1] = {hdr.c_str()};
  const char* hn[1] = {"vx_field.h"};
  hiprtcProgram prog;
  rtc_check(hiprtcCreateProgram(&prog, src.c_str(), "bc_jit.hip", 1, hs, hn), "create");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, 6, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw Error(ERR_INTERNAL, "bytecode JIT: compile failed: " + log.substr(0, 2000));
  }
  size_t n = 0;
  rtc_check(hiprtcGetCodeSize(prog, &n), "code size");
  std::string code(n, '\0');
  rtc_check(hiprtcGetCode(prog, &code[0]), "code");
  hiprtcDestroyProgram(&prog);
  return code;
}

JitKernel compile(const std::string& src) {
  const std::string code = compile_code(src);
  JitKernel k;
  VX_HIP(hipModuleLoadData(&k.mod, code.data()));
  VX_HIP(hipModuleGetFunction(&k.fn, k.mod, "bc_jit"));
  return k;
}

}  // namespace

size_t bytecode_jit_code_bytes(const uint32_t* words, uint32_t n_words, uint32_t n_trace_cols, uint32_t eval_log,
                               uint32_t trace_log) {
  size_t n_ops = 0;
  return compile_code(bc_jit_source(words, n_words, n_trace_cols, eval_log, trace_log, &n_ops)).size();
}

void bytecode_jit_enable(bool on) {
  std::lock_guard<std::mutex> g(g_mu);
  g_enabled = on;
}
bool bytecode_jit_enabled() {
  std::lock_guard<std::mutex> g(g_mu);
  return g_enabled;
}

void bytecode_jit_eval(const uint32_t* words, uint32_t n_words, const uint32_t* const* d_cols, uint32_t n_trace_cols,
                       uint32_t n_rows, uint32_t trace_n_rows, const uint32_t* coeff, const uint32_t* denom_inv,
                       uint32_t* const acc[4], hipStream_t s) {
  const uint32_t e = 31u - (uint32_t)__builtin_clz(n_rows), d = 31u - (uint32_t)__builtin_clz(trace_n_rows);
  int dev = 0;
  VX_HIP(hipGetDevice(&dev));
  std::string key(reinterpret_cast<const char*>(words), 4ull * n_words);
  key += "|" + std::to_string(dev) + "|" + std::to_string(e) + "|" + std::to_string(d) + "|" +
         std::to_string(n_trace_cols);
  JitKernel k;
  {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_kernels.find(key);
    if (it == g_kernels.end()) {
      size_t n_ops = 0;
      const std::string src = bc_jit_source(words, n_words, n_trace_cols, e, d, &n_ops);
      it = g_kernels.emplace(key, compile(src)).first;
    }
    k = it->second;
  }
  uint32_t* a0 = acc[0];
  uint32_t* a1 = acc[1];
  uint32_t* a2 = acc[2];
  uint32_t* a3 = acc[3];
  void* args[] = {(void*)&d_cols, (void*)&coeff, (void*)&denom_inv, &a0, &a1, &a2, &a3, &n_rows};
  VX_HIP(hipModuleLaunchKernel(k.fn, (n_rows + 255) / 256, 1, 1, 256, 1, 1, 0, s, args, nullptr));
}

}  // namespace vx
