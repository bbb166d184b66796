// Run-time compiled stage kernels.
//
// The stage kernels of tqsv_stage.hip interpret a stage program: every op is
// decoded from its words and dispatched through a switch on (kind, target
// register bit, flags).  On gfx950 that costs about 40% extra vector
// instructions over the gate arithmetic (loop-carried register copies at
// the switch joins, run-time control selects), and those kernels are bound
// by vector issue, not by HBM.  The host planner can instead emit, for one
// stage program, HIP source in which every op is straight-line code with its
// bit positions, controls and matrix offsets as constants; hipRTC compiles it
// here for gfx950 and the launchers below run it with the interpreter's
// launch shape and argument meaning.
//
// Entry points (see tqsv.h): tqsv_jit_compile (source -> code object, no GPU
// needed), tqsv_jit_load (code object -> module + kernel handles),
// tqsv_jit_stage_forward / tqsv_jit_stage_adjoint (launch one stage).
#include <hip/hiprtc.h>

#include <string>
#include <vector>

#include "tqsv_common.h"

int stage_launch_shape(const int* prog_host, int n, int64_t bsz, int grid_cap, int& T, int64_t& n_groups,
                       int& grid, int64_t& gpw, size_t& shm);
int64_t stage_rows_per_wg(const int* prog_host, int64_t gpw);
int64_t stage_adjoint_work(const int* prog_host, int64_t gpw, int grid);
int stage_adjoint_reduce(hipStream_t s, const int* prog_host, const int* prog_dev, int64_t bsz, int64_t gpw,
                         int grid, const float* work, float* grad, int64_t grad_stride);
using namespace tqsv;

namespace {


constexpr int kHdrT = 2;    // stage header word: thread bits T (2^T threads per group)
constexpr int kHdrNgp = 5;  // stage header word: number of gradient parameters

// gfx950 code, no packed-FP32 instructions (the same flags as the library)
const char* const kOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                             "-Xclang", "-packed-fp32-ops"};

}  // namespace

// Compile HIP source to a gfx950 code object.  opts (optional): extra
// compiler options, separated by spaces (e.g. an LLVM scheduling strategy);
// the library's own flags always apply.  code/cap: output buffer (may be
// null to query the size); *size receives the code object's size; log
// (optional) receives the compiler log, truncated to log_cap - 1 characters.
TQSV_API int tqsv_jit_compile(const char* src, const char* opts, void* code, int64_t cap, int64_t* size, char* log,
                              int64_t log_cap) {
  if (!src || !size) { set_error("tqsv_jit_compile: null argument"); return E_ARG; }
  if (log && log_cap > 0) log[0] = 0;
  std::vector<std::string> extra;
  if (opts) {
    std::string cur;
    for (const char* c = opts; ; ++c) {
      if (*c == ' ' || *c == 0) {
        if (!cur.empty()) extra.push_back(cur);
        cur.clear();
        if (*c == 0) break;
      } else {
        cur += *c;
      }
    }
  }
  std::vector<const char*> all(kOpts, kOpts + sizeof(kOpts) / sizeof(kOpts[0]));
  for (const std::string& e : extra) all.push_back(e.c_str());
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, src, "tqsv_stage_jit.hip", 0, nullptr, nullptr);
  if (r != HIPRTC_SUCCESS) { set_error("tqsv_jit_compile: hiprtcCreateProgram: %s", hiprtcGetErrorString(r)); return E_LAUNCH; }
  r = hiprtcCompileProgram(prog, (int)all.size(), all.data());
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  std::string text(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &text[0]);
  if (log && log_cap > 0) {
    const size_t k = text.size() < (size_t)(log_cap - 1) ? text.size() : (size_t)(log_cap - 1);
    text.copy(log, k);
    log[k] = 0;
  }
  if (r != HIPRTC_SUCCESS) {
    set_error("tqsv_jit_compile: %s: %.400s", hiprtcGetErrorString(r), text.c_str());
    hiprtcDestroyProgram(&prog);
    return E_ARG;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  *size = (int64_t)cs;
  if (code) {
    if (cap < (int64_t)cs) {
      hiprtcDestroyProgram(&prog);
      set_error("tqsv_jit_compile: buffer of %lld bytes < code object of %zu", (long long)cap, cs);
      return E_ARG;
    }
    hiprtcGetCode(prog, (char*)code);
  }
  hiprtcDestroyProgram(&prog);
  return OK;
}

// Load a code object on the current device and look up n kernels by name.
// The module stays loaded until tqsv_jit_unload.
TQSV_API int tqsv_jit_load(const void* code, int64_t size, int n, const char* const* names, void** module,
                           void** funcs) {
  if (!code || size <= 0 || n < 0 || (n > 0 && (!names || !funcs)) || !module) {
    set_error("tqsv_jit_load: bad argument");
    return E_ARG;
  }
  hipModule_t mod = nullptr;
  hipError_t e = hipModuleLoadData(&mod, code);
  if (e != hipSuccess) { set_error("tqsv_jit_load: hipModuleLoadData: %s", hipGetErrorString(e)); return E_LAUNCH; }
  for (int i = 0; i < n; ++i) {
    hipFunction_t f = nullptr;
    e = hipModuleGetFunction(&f, mod, names[i]);
    if (e != hipSuccess) {
      set_error("tqsv_jit_load: kernel '%s': %s", names[i], hipGetErrorString(e));
      hipModuleUnload(mod);
      return E_ARG;
    }
    funcs[i] = (void*)f;
  }
  *module = (void*)mod;
  return OK;
}

TQSV_API int tqsv_jit_unload(void* module) {
  if (!module) return OK;
  hipError_t e = hipModuleUnload((hipModule_t)module);
  if (e != hipSuccess) { set_error("tqsv_jit_unload: %s", hipGetErrorString(e)); return E_LAUNCH; }
  return OK;
}

// One stage of a forward pass with a compiled stage kernel.  Arguments as
// tqsv_stage_forward; func: the kernel compiled from the same program
// (kernel parameters: out, in, mats, init_zero, n_groups, groups_per_wg).
TQSV_API int tqsv_jit_stage_forward(hipStream_t s, const void* func, void* state_out, const void* state_in,
                                    int64_t bsz, int n, const int* prog_host, const void* mats, int init_zero,
                                    int grid_cap) {
  if (!func || !state_out || !prog_host || !mats || (!state_in && !init_zero)) {
    set_error("tqsv_jit_stage_forward: null pointer argument");
    return E_ARG;
  }
  if (bsz < 1 || n < 4 || n > 32) { set_error("tqsv_jit_stage_forward: bad bsz/n"); return E_ARG; }
  int T, grid;
  int64_t n_groups, gpw;
  size_t shm;
  const int Th = prog_host[kHdrT];
  if (Th < 6 || Th > 9) { set_error("tqsv_jit_stage_forward: compiled stages need 64..512 threads (6 <= T <= 9)"); return E_ARG; }
  int rc = stage_launch_shape(prog_host, n, bsz, grid_cap, T, n_groups, grid, gpw, shm);
  if (rc) return rc;
  void* out = state_out;
  const void* in = state_in;
  const void* m = mats;
  int iz = init_zero;
  long long ng = n_groups, gp = gpw;
  void* args[] = {&out, &in, &m, &iz, &ng, &gp};
  hipError_t e = hipModuleLaunchKernel((hipFunction_t)func, grid, 1, 1, 1u << T, 1, 1, 0, s, args, nullptr);
  if (e != hipSuccess) { set_error("tqsv_jit_stage_forward: %s", hipGetErrorString(e)); return E_LAUNCH; }
  return OK;
}

// One stage of the adjoint sweep with a compiled stage kernel.  Arguments as
// tqsv_stage_adjoint, plus psi_in / lam_in: where psi and lam are read (NULL =
// in place); the swept states are written to psi / lam.  The kernel writes
// per-workgroup gradient partials to `work`, which a fixed-order reduction
// then adds into grad (its gslot table read from prog_dev).  Kernel
// parameters: psi, lam, psi_in, lam_in, mats, partials, rows per workgroup,
// n_groups, groups_per_wg.
TQSV_API int tqsv_jit_stage_adjoint(hipStream_t s, const void* func, void* psi, void* lam, const void* psi_in,
                                    const void* lam_in, int64_t bsz, int n, const int* prog_host,
                                    const int* prog_dev, const void* mats, float* grad, int64_t grad_stride,
                                    float* work, int64_t work_floats, int grid_cap) {
  if (!func || !psi || !lam || !prog_host || !prog_dev || !mats) {
    set_error("tqsv_jit_stage_adjoint: null pointer argument");
    return E_ARG;
  }
  if (bsz < 1 || n < 4 || n > 32) { set_error("tqsv_jit_stage_adjoint: bad bsz/n"); return E_ARG; }
  if (prog_host[kHdrNgp] > 0 && (!grad || !work)) {
    set_error("tqsv_jit_stage_adjoint: grad and partial buffers required");
    return E_ARG;
  }
  int T, grid;
  int64_t n_groups, gpw;
  size_t shm;
  const int Th = prog_host[kHdrT];
  if (Th < 6 || Th > 9) { set_error("tqsv_jit_stage_adjoint: compiled stages need 64..512 threads (6 <= T <= 9)"); return E_ARG; }
  int rc = stage_launch_shape(prog_host, n, bsz, grid_cap, T, n_groups, grid, gpw, shm);
  if (rc) return rc;
  const int64_t need = stage_adjoint_work(prog_host, gpw, grid);
  if (need > work_floats) {
    set_error("tqsv_jit_stage_adjoint: partial buffer of %lld floats, the stage needs %lld "
              "(tqsv_stage_adjoint_work_floats)", (long long)work_floats, (long long)need);
    return E_ARG;
  }
  void* p = psi;
  void* l = lam;
  const void* pin = psi_in ? psi_in : psi;
  const void* lin = lam_in ? lam_in : lam;
  const void* m = mats;
  float* part = work;
  long long rpw = stage_rows_per_wg(prog_host, gpw), ng = n_groups, gp = gpw;
  void* args[] = {&p, &l, &pin, &lin, &m, &part, &rpw, &ng, &gp};
  hipError_t e = hipModuleLaunchKernel((hipFunction_t)func, grid, 1, 1, 1u << T, 1, 1, 0, s, args, nullptr);
  if (e != hipSuccess) { set_error("tqsv_jit_stage_adjoint: %s", hipGetErrorString(e)); return E_LAUNCH; }
  stage_adjoint_reduce(s, prog_host, prog_dev, bsz, gpw, grid, work, grad, grad_stride);
  e = hipGetLastError();
  if (e != hipSuccess) { set_error("tqsv_jit_stage_adjoint: %s", hipGetErrorString(e)); return E_LAUNCH; }
  return OK;
}
