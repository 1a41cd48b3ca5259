// bmx_jit.cpp -- runtime-compiled cell kernels.
//
// The element loop of the kernels is an interpreter: per op it reads the op's
// type and flags and branches to the op's code.  Real lattices repeat a cell
// (a fixed sequence of element kinds with different strengths), so the op
// program is mostly one window of op types and flags, repeated.  For such a
// lattice this file finds that window (the cell), generates jit_cell(): the
// cell's ops as calls of run_op with each op's type, flags and count as
// compile-time constants (so every dispatch test folds away, kick orders and
// slice loops are unrolled, and the compiler schedules across op
// boundaries), and compiles bmx_ops.h + jit_cell + bmx_kernels.h
// with hipRTC for the device.  The walk then tracks every run of the cell
// through jit_cell and the remaining ops through run_op.  Both evaluate the
// same op functions in the same order.
//
// Compiled modules are kept per (device, cell) for the life of the process.
// BMX_JIT=0 in the environment turns the cell kernels off; BMX_JIT_FLAGS adds
// compiler options (e.g. -DBMX_WAVES_PER_EU=3).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bmx.h"
#include "bmx_internal.h"
#include "bmx_jit.h"

namespace bmx {

namespace {

// header texts, generated from the sources by the build (bmx_jit_src.inc)
#include "bmx_jit_src.inc"

// type (8 bits), flags (24 bits) and n (32 bits): what the cell kernel's
// code is specialised on
uint64_t op_sig(const DevOp& op) {
  return ((uint64_t)(uint32_t)op.n << 32) | ((uint64_t)((uint32_t)op.flags & ~(uint32_t)F_CELL & 0xffffffu) << 8) |
         (uint32_t)(op.type & 0xff);
}

bool ele_start(const std::vector<DevOp>& prog, size_t k) { return k == 0 || prog[k].ie != prog[k - 1].ie; }

// Most frequent window of `len` ops starting at an element boundary among
// the first n ops; returns its start (-1 if none) and its count.
int64_t frequent_window(const std::vector<DevOp>& prog, size_t n, int len, int* count) {
  const uint64_t B = 0x100000001b3ull;
  std::vector<uint64_t> sig(n);
  for (size_t i = 0; i < n; i++) sig[i] = op_sig(prog[i]) * 0x9e3779b97f4a7c15ull + 1;
  uint64_t pw = 1;
  for (int i = 0; i < len; i++) pw *= B;
  std::unordered_map<uint64_t, std::pair<int, int64_t>> seen;  // hash -> (count, first start)
  uint64_t h = 0;
  int64_t best = -1;
  *count = 0;
  for (size_t k = 0; k < n; k++) {
    h = h * B + sig[k];
    if (k >= (size_t)len) h -= pw * sig[k - len];
    if (k + 1 < (size_t)len) continue;
    size_t start = k + 1 - len;
    if (!ele_start(prog, start)) continue;
    auto& e = seen[h];
    if (e.first == 0) e.second = (int64_t)start;
    if (++e.first > *count) {
      *count = e.first;
      best = e.second;
    }
  }
  return best;
}

bool same_window(const std::vector<DevOp>& prog, size_t a, size_t b, int len) {
  for (int i = 0; i < len; i++)
    if (op_sig(prog[a + i]) != op_sig(prog[b + i])) return false;
  return true;
}

// Greedy parse: cell starts of non-overlapping matches of the window at rep,
// over the first `limit` ops (all when 0).
std::vector<size_t> parse_cells(const std::vector<DevOp>& prog, size_t rep, int len, size_t limit = 0) {
  std::vector<size_t> starts;
  size_t n = (limit && limit < prog.size()) ? limit : prog.size();
  for (size_t k = 0; k + len <= n;) {
    if (ele_start(prog, k) && same_window(prog, k, rep, len)) {
      starts.push_back(k);
      k += len;
    } else {
      k++;
    }
  }
  return starts;
}

struct Compiled {
  hipModule_t mod;
  JitKernels k;
};

std::mutex g_mu;
std::map<std::pair<int, std::string>, Compiled> g_cache;

bool jit_enabled() {
  const char* e = std::getenv("BMX_JIT");
  return !(e && e[0] == '0');
}

// absolute-time rf anywhere in the program: the lanes carry their time offset
bool has_abs_time(const std::vector<DevOp>& prog) {
  for (const DevOp& op : prog)
    if (op.type == OP_RFCAVITY && (op.flags & F_RF_ABS)) return true;
  return false;
}

// radiation steps anywhere in the program: the lanes carry their particle index
bool has_rad(const std::vector<DevOp>& prog) {
  for (const DevOp& op : prog)
    if (op.type == OP_RAD) return true;
  return false;
}

std::string cell_source(const std::vector<DevOp>& prog, size_t rep, int len) {
  std::string s;
  s += "#include \"bmx.h\"\n#include \"bmx_internal.h\"\nusing namespace bmx;\n";
  s += "#define BMX_CELL_LEN " + std::to_string(len) + "\n";
  s += std::string("#define BMX_ABS_TIME ") + (has_abs_time(prog) ? "1" : "0") + "\n";
  s += std::string("#define BMX_RAD ") + (has_rad(prog) ? "1" : "0") + "\n";
  s += kSrcOpsH;
  s += "\ntemplate <class Loss>\n__device__ __forceinline__ void jit_cell(Lane& L, const DevOp* o, "
       "const LatticeConsts& lc, const TurnTime& tt, Loss loss) {\n";
  for (int i = 0; i < len; i++) {
    const DevOp& op = prog[rep + i];
    s += "  run_op(L, o[" + std::to_string(i) + "], " + std::to_string(op.type) + ", " +
         std::to_string(op.flags & ~F_CELL) + ", " + std::to_string(op.n) + ", lc, tt);\n  loss(o[" +
         std::to_string(i) + "]);\n";
  }
  s += "}\n";
  s += kSrcKernelsH;
  return s;
}

}  // namespace

bool plan_cells(std::vector<DevOp>& prog, CellPlan* plan) {
  plan->len = 0;
  plan->key.clear();
  plan->cell_prv.clear();
  size_t n = prog.size();
  if (!jit_enabled() || n < 64) return false;
  size_t n_probe = n < 20000 ? n : 20000;
  int best_len = 0;
  size_t best_cov = 0, best_rep = 0;
  for (int len = 2; len <= 64 && (size_t)len * 4 <= n_probe; len++) {
    int count = 0;
    int64_t rep = frequent_window(prog, n_probe, len, &count);
    if (rep < 0 || count < 4) continue;
    // coverage over the probe prefix: a multi-variant program repeats its
    // variants, so the prefix ranks the window lengths as the whole does
    size_t cov = parse_cells(prog, (size_t)rep, len, n_probe).size() * (size_t)len;
    // a longer cell must cover clearly more ops than a shorter one
    if (cov > best_cov + best_cov / 50) {
      best_cov = cov;
      best_len = len;
      best_rep = (size_t)rep;
    }
  }
  if (best_len == 0 || best_cov * 2 < n_probe) return false;
  std::vector<size_t> starts = parse_cells(prog, best_rep, best_len);
  if (starts.size() * (size_t)best_len * 2 < n) return false;
  for (auto& op : prog) op.flags &= ~F_CELL;
  for (size_t k : starts) prog[k].flags |= F_CELL;
  plan->cell_prv.assign(n, -1);
  int32_t last = -1;
  size_t j = 0;
  for (size_t i = 0; i < n; i++) {
    if (j < starts.size() && starts[j] == i) {
      last = (int32_t)i;
      j++;
    }
    plan->cell_prv[i] = last;
  }
  plan->len = best_len;
  plan->rep = best_rep;
  plan->n_cells = starts.size();
  for (int i = 0; i < best_len; i++) {
    const DevOp& op = prog[best_rep + i];
    plan->key += std::to_string(op.type) + ":" + std::to_string(op.flags & ~F_CELL) + ":" + std::to_string(op.n) + ";";
  }
  return true;
}

const JitKernels* jit_kernels(int device, const std::vector<DevOp>& prog, const CellPlan& plan, std::string* err) {
  std::lock_guard<std::mutex> lock(g_mu);
  const char* fl = std::getenv("BMX_JIT_FLAGS");
  auto key = std::make_pair(device, plan.key + "|" + (has_abs_time(prog) ? "abs|" : "") + (has_rad(prog) ? "rad|" : "") +
                                        (fl ? fl : ""));
  auto it = g_cache.find(key);
  if (it != g_cache.end()) return &it->second.k;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    *err = "hipGetDeviceProperties failed";
    return nullptr;
  }
  std::string arch = prop.gcnArchName;
  std::string src = cell_source(prog, plan.rep, plan.len);
  const char* headers[] = {kSrcBmxH, kSrcInternalH};
  const char* names[] = {"bmx.h", "bmx_internal.h"};
  hiprtcProgram p;
  if (hiprtcCreateProgram(&p, src.c_str(), "bmx_cell.hip", 2, headers, names) != HIPRTC_SUCCESS) {
    *err = "hiprtcCreateProgram failed";
    return nullptr;
  }
  std::string arch_opt = "--offload-arch=" + arch;
  std::vector<std::string> extra;  // BMX_JIT_FLAGS: further compiler options, space separated
  if (const char* e = std::getenv("BMX_JIT_FLAGS")) {
    std::string f(e), w;
    for (char ch : f + " ") {
      if (ch == ' ') { if (!w.empty()) extra.push_back(w); w.clear(); } else w += ch;
    }
  }
  // the launch geometry the host uses for these kernels
  std::string np_opt = "-DBMX_NP=" + std::to_string(BMX_NP);
  std::string gb_opt = "-DBMX_GRID_BLOCK=" + std::to_string(BMX_GRID_BLOCK);
  std::vector<const char*> opts = {arch_opt.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                                   "-packed-fp32-ops", np_opt.c_str(), gb_opt.c_str()};
  for (auto& x : extra) opts.push_back(x.c_str());
  hiprtcResult r = hiprtcCompileProgram(p, (int)opts.size(), opts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(p, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(p, &log[0]);
    *err = std::string("cell kernel compile failed: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 2000);
    hiprtcDestroyProgram(&p);
    return nullptr;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(p, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(p, code.data());
  hiprtcDestroyProgram(&p);
  Compiled c;
  if (hipModuleLoadData(&c.mod, code.data()) != hipSuccess) {
    *err = "hipModuleLoadData of the cell kernels failed";
    return nullptr;
  }
  if (hipModuleGetFunction(&c.k.bunch, c.mod, "bmx_track_bunch_kernel") != hipSuccess ||
      hipModuleGetFunction(&c.k.multi, c.mod, "bmx_track_multi_kernel") != hipSuccess ||
      hipModuleGetFunction(&c.k.grid, c.mod, "bmx_grid_block_kernel") != hipSuccess) {
    *err = "cell kernels missing from the compiled module";
    return nullptr;
  }
  return &g_cache.emplace(key, c).first->second.k;
}

}  // namespace bmx
